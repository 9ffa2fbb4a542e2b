"""Contract of the drop-in Python surface, re-expressed from the reference's
own tests (python/test/test_device/test_gpu/test_libepseon_gpu.py and
test_cpu/test_libepseon_cpu.py).

The parts that need a device are marked ``gpu``; everything else (builder
identity, precision parsing, error messages, C ABI exports) runs on the CPU.
"""
from __future__ import annotations

import ctypes
import gc
import re
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]
HEADER = REPO / "include" / "epseon_hip.h"

COMPUTE_GROUP_AXES_COUNT = 3


def _gpu():
    from epseon_backend_amd.device.gpu import _libepseon_gpu

    return _libepseon_gpu


def _morse(m, n=16500):
    return m.MorsePotentialConfig(
        dissociation_energy=5500.0,
        equilibrium_bond_distance=0.6,
        well_width=10,
        min_r=0.0,
        max_r=10.0,
        point_count=n,
    )


def _configure(m, configurator):
    """The reference test's canonical builder chain (test_libepseon_gpu.py:172-209)."""
    return (
        configurator.set_hardware_config(
            potential_buffer_size=16500,
            group_size=512,
            allocation_block_size=16 * 1024 * 1024,
        )
        .set_morse_potential([_morse(m), _morse(m)])
        .set_vibwa_algorithm(
            mass_atom_0=87.62,
            mass_atom_1=87.62,
            integration_step=0.1,
            min_distance_to_asymptote=0.1,
            min_level=0,
            max_level=0,
        )
    )


# ---------------------------------------------------------------- CPU side


def test_greet() -> None:
    from epseon_backend_amd.device.cpu._libepseon_cpu import greet

    assert greet() == "Hello, World from C++!"


def test_context_creates_without_device_error() -> None:
    m = _gpu()
    ctx = m.EpseonComputeContext.create()
    assert repr(ctx)
    assert re.match(r"\d+\.\d+\.\d+\.\d+", ctx.get_vulkan_version()) is not None


@pytest.mark.parametrize("precision", ["float32", "float64", "FLOAT32", "Float64"])
def test_configurator_builder_returns_same_object(precision) -> None:
    m = _gpu()
    cls = m.TaskConfiguratorFloat32 if precision.lower() == "float32" else m.TaskConfiguratorFloat64
    c = cls()
    a = c.set_hardware_config(potential_buffer_size=16500, group_size=4096, allocation_block_size=64 << 20)
    b = a.set_morse_potential([_morse(m), _morse(m)])
    d = b.set_vibwa_algorithm(
        mass_atom_0=87.62,
        mass_atom_1=87.62,
        integration_step=0.1,
        min_distance_to_asymptote=0.1,
        min_level=0,
        max_level=0,
    )
    assert id(c) == id(a) == id(b) == id(d)
    assert d.is_configured()


def test_configurator_is_configured_stages() -> None:
    m = _gpu()
    c = m.TaskConfiguratorFloat64()
    assert not c.is_configured()
    c.set_hardware_config(potential_buffer_size=10, group_size=1, allocation_block_size=1 << 20)
    assert not c.is_configured()
    c.set_morse_potential([_morse(m, 10)])
    assert not c.is_configured()
    c.set_vibwa_algorithm(
        mass_atom_0=1.0, mass_atom_1=1.0, integration_step=0.1, min_distance_to_asymptote=0.0,
        min_level=0, max_level=1,
    )
    assert c.is_configured()


def test_mismatched_point_count_message() -> None:
    m = _gpu()
    c = m.TaskConfiguratorFloat32()
    with pytest.raises(RuntimeError, match="All Morse potentials must have same point count, but previous ones had 16500"):
        c.set_morse_potential([_morse(m, 16500), _morse(m, 16400)])
    # The batch setter (an addition) accepts ragged grids.
    c.set_morse_potential_batch([_morse(m, 16500), _morse(m, 16400)])


def test_morse_config_fields() -> None:
    m = _gpu()
    c = _morse(m)
    assert (c.dissociation_energy, c.equilibrium_bond_distance, c.well_width) == (5500.0, 0.6, 10.0)
    assert (c.min_r, c.max_r, c.point_count) == (0.0, 10.0, 16500)


def test_invalid_options_raise() -> None:
    m = _gpu()
    c = m.TaskConfiguratorFloat64()
    with pytest.raises(RuntimeError):
        c.set_trial_energies(0)
    with pytest.raises(RuntimeError):
        c.set_reduced_masses([1.0, -2.0])
    with pytest.raises(RuntimeError):
        c.set_tolerance(-1.0)
    with pytest.raises(RuntimeError):
        m.DoubleWellPotentialConfig(1.0, 1.0, 1.0, 1.0, 1.0, 0.0, 0.0, 1.0, 100)
    with pytest.raises(RuntimeError):
        c.set_tabulated_potential([m.TabulatedPotentialConfig([1.0, 2.0], 0.0, 1.0, 3.0)])


def test_no_device_interface_without_device() -> None:
    m = _gpu()
    ctx = m.EpseonComputeContext.create()
    n = len(list(ctx.get_physical_device_info()))
    with pytest.raises(RuntimeError, match="Device not available."):
        ctx.get_device_interface(n)


def _declared_functions() -> list[str]:
    text = HEADER.read_text()
    return re.findall(r"^(?:int|const char\*)\s+(epseon_\w+)\(", text, flags=re.M)


def test_capi_exports_every_declared_symbol() -> None:
    from epseon_backend_amd import capi

    lib = capi.load_library()
    names = _declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
        assert name in capi.SIGNATURES, f"{name} missing from the ctypes binding"
    # and the binding declares nothing the header does not
    assert set(capi.SIGNATURES) == set(names)


def test_capi_errors_without_device() -> None:
    from epseon_backend_amd import capi

    if capi.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(capi.EpseonError, match="Device not available"):
        capi.Context(0)
    lib = capi.load_library()
    out = ctypes.c_void_p()
    assert lib.epseon_ctx_create(0, 0, ctypes.byref(out)) == capi.EPSEON_ERR_DEVICE
    assert lib.epseon_scan(None) == capi.EPSEON_ERR_INVALID
    assert b"null context" in lib.epseon_last_error()


# ---------------------------------------------------------------- GPU side


@pytest.mark.gpu
def test_get_physical_device_info_device_properties() -> None:
    m = _gpu()
    ctx = m.EpseonComputeContext.create()
    infos = list(ctx.get_physical_device_info())
    assert len(infos) > 0
    for device in infos:
        p = device.device_properties
        assert isinstance(p.api_version, str)
        assert isinstance(p.driver_version, str)
        assert isinstance(p.vendor_id, int) and p.vendor_id == 0x1002
        assert isinstance(p.device_id, int)
        assert p.device_type in ("INTEGRATED_GPU", "DISCRETE_GPU", "VIRTUAL_GPU", "CPU", "OTHER")
        assert isinstance(p.device_name, str) and len(p.device_name) > 0
        assert isinstance(p.pipeline_cache_uuid, list)
        assert isinstance(p.limits.max_compute_shared_memory_size, int)
        assert isinstance(p.limits.max_compute_work_group_count, tuple)
        assert len(p.limits.max_compute_work_group_count) == COMPUTE_GROUP_AXES_COUNT
        assert isinstance(p.limits.max_compute_work_group_invocations, int)
        assert isinstance(p.limits.max_compute_work_group_size, int)
        assert hasattr(p, "sparse_properties")
        assert p.arch.startswith("gfx950")


@pytest.mark.gpu
def test_get_physical_device_info_memory_properties() -> None:
    m = _gpu()
    ctx = m.EpseonComputeContext.create()
    for device in ctx.get_physical_device_info():
        heaps = device.memory_properties.memory_heaps
        assert heaps and heaps[0].size > 100 * 2**30
        assert "DEVICE_LOCAL" in heaps[0].flags
        assert device.memory_properties.memory_types[0].heap_index == 0


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_get_and_use_task_configurator(precision) -> None:
    m = _gpu()
    ctx = m.EpseonComputeContext.create()
    info = next(iter(ctx.get_physical_device_info()))
    interface = ctx.get_device_interface(info.device_properties.device_id)
    cfg = _configure(m, interface.get_task_configurator(precision))
    assert id(cfg)


@pytest.mark.gpu
def test_unknown_precision() -> None:
    m = _gpu()
    ctx = m.EpseonComputeContext.create()
    interface = ctx.get_device_interface(0)
    with pytest.raises(ValueError, match='Invalid PrecisionType literal in string: "float80"'):
        interface.get_task_configurator("float80")


@pytest.mark.gpu
def test_submit_unconfigured_raises() -> None:
    m = _gpu()
    interface = m.EpseonComputeContext.create().get_device_interface(0)
    c = interface.get_task_configurator("float64")
    with pytest.raises(RuntimeError, match="TaskConfigurator submitted for execution before fully configured."):
        interface.submit_task(c)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_submit_task_with_ctx_going_out_of_scope(precision) -> None:
    m = _gpu()
    ctx = m.EpseonComputeContext.create()
    interface = ctx.get_device_interface(0)
    del ctx
    gc.collect(0)
    gc.collect(1)
    gc.collect(2)
    cfg = _configure(m, interface.get_task_configurator(precision))
    handle = interface.submit_task(cfg)
    assert isinstance(handle.get_status_message(), str)
    handle.wait()
    assert handle.is_done()
    assert handle.get_status_message() == "done"
    levels = handle.get_levels()
    assert levels.shape == (2, 1)
    assert levels.dtype == (np.float32 if precision == "float32" else np.float64)
    # Heavy Sr2-like Morse: the ground level is below the first excited one
    # and close to the harmonic estimate.
    assert np.all(handle.get_level_status() == 0)
    mu = 87.62 / 2
    kp = 16.857629191640175 / mu
    e0 = 2 * 10 * np.sqrt(kp * 5500.0) * 0.5 - 100 * kp * 0.25
    np.testing.assert_allclose(levels[:, 0], e0, rtol=1e-4)


@pytest.mark.gpu
def test_submit_and_wait_every_device() -> None:
    m = _gpu()
    ctx = m.EpseonComputeContext.create()
    for info in ctx.get_physical_device_info():
        interface = ctx.get_device_interface(info.device_properties.device_id)
        handle = interface.submit_task(_configure(m, interface.get_task_configurator("float32")))
        handle.wait()
        assert handle.is_done()


@pytest.mark.gpu
def test_builder_snapshot_at_submit() -> None:
    """Changing the builder after submit does not change the running task."""
    m = _gpu()
    interface = m.EpseonComputeContext.create().get_device_interface(0)
    c = _configure(m, interface.get_task_configurator("float64"))
    handle = interface.submit_task(c)
    c.set_vibwa_algorithm(
        mass_atom_0=1.0, mass_atom_1=1.0, integration_step=0.1, min_distance_to_asymptote=0.1,
        min_level=0, max_level=5,
    )
    handle.wait()
    assert handle.get_levels().shape == (2, 1)


@pytest.mark.gpu
def test_worker_error_is_raised_on_wait() -> None:
    m = _gpu()
    interface = m.EpseonComputeContext.create().get_device_interface(0)
    c = _configure(m, interface.get_task_configurator("float64"))
    c.set_hardware_config(potential_buffer_size=100, group_size=1, allocation_block_size=1 << 20)
    handle = interface.submit_task(c)
    with pytest.raises(RuntimeError, match="exceeds potential_buffer_size"):
        handle.wait()
    assert handle.is_done()
    assert handle.get_status_message().startswith("failed")


@pytest.mark.gpu
def test_eigenfunction_output_through_handle() -> None:
    """Additive: set_eigenfunction_output returns the builder and the handle
    then carries normalised eigenfunctions (sum psi^2 h = 1), one
    (levels, points) array per potential; without it the getter raises."""
    m = _gpu()
    interface = m.EpseonComputeContext.create().get_device_interface(0)
    c = _configure(m, interface.get_task_configurator("float64"))
    c.set_vibwa_algorithm(mass_atom_0=87.62, mass_atom_1=87.62, integration_step=0.1,
                          min_distance_to_asymptote=0.1, min_level=0, max_level=3)
    assert c.set_eigenfunction_output(True) is c
    handle = interface.submit_task(c)
    handle.wait()
    psi = handle.get_wavefunctions()
    assert len(psi) == 2 and psi[0].shape == (4, 16500)
    h = 10.0 / 16499
    np.testing.assert_allclose((psi[0] ** 2).sum(axis=1) * h, 1.0, rtol=1e-12)
    # Level v has v nodes inside the well.
    for v in range(4):
        row = psi[0][v]
        inner = row[np.abs(row) > 1e-6 * np.abs(row).max()]
        assert np.sum(np.diff(np.sign(inner)) != 0) == v
    plain = interface.submit_task(_configure(m, interface.get_task_configurator("float64")))
    plain.wait()
    with pytest.raises(RuntimeError, match="set_eigenfunction_output"):
        plain.get_wavefunctions()


@pytest.mark.gpu
def test_drop_in_path_matches_oracle() -> None:
    """The H2 workload through the pybind11 drop-in path (configurator,
    submit, wait, handle getters) equals the oracle run on the same Morse
    tabulation: node counts, far-end values, levels and iterations
    bit-exact."""
    import oracle

    m = _gpu()
    interface = m.EpseonComputeContext.create().get_device_interface(0)
    de, re, a, r_min, r_max, n = 38297.0, 0.7414, 1.9426, 0.2, 5.0, 10000
    mass = 2 * 0.50391
    c = (
        interface.get_task_configurator("float64")
        .set_hardware_config(potential_buffer_size=n, group_size=1, allocation_block_size=16 << 20)
        .set_morse_potential([m.MorsePotentialConfig(dissociation_energy=de, equilibrium_bond_distance=re,
                                                     well_width=a, min_r=r_min, max_r=r_max, point_count=n)])
        .set_vibwa_algorithm(mass_atom_0=mass, mass_atom_1=mass, integration_step=0.1,
                             min_distance_to_asymptote=1.0, min_level=0, max_level=16)
    )
    handle = interface.submit_task(c)
    handle.wait()
    levels = handle.get_levels()[0]
    nodes = handle.get_node_counts()[0]
    far = handle.get_far_end_values()[0]
    its = handle.get_iterations()[0]
    status = handle.get_level_status()[0]
    # The worker's defaults: 320 trial energies for 17 levels (16 per level,
    # whole 64-energy groups, at least 256), whole-grid shots (one small
    # potential), segmented refinement in 256-lane blocks.
    mu = mass * mass / (mass + mass)
    v = oracle.morse(de, re, a, r_min, oracle.grid_step(r_min, r_max, n), n)
    prob = oracle.Problem(v=v, r_min=r_min, r_max=r_max, mu=mu, e_top=de - 1.0, m=320)
    on, of, orr, olv, oit = oracle.solve(prob, 0, 17, 1e-10, 200, segmented=True, lanes=256, windows=False)
    np.testing.assert_array_equal(nodes, on)
    np.testing.assert_array_equal(far.view(np.int64), of.view(np.int64))
    # Levels the window does not bracket (v = 16 lies within 1 cm^-1 of D_e)
    # are NaN with status UNBOUND on both sides.
    ok = ~np.isnan(olv)
    assert ok.sum() >= 16
    np.testing.assert_array_equal(status[ok], 0)
    np.testing.assert_array_equal(status[~ok], 1)
    np.testing.assert_array_equal(np.isnan(levels), ~ok)
    np.testing.assert_array_equal(levels[ok].view(np.int64), olv[ok].view(np.int64))
    np.testing.assert_array_equal(its[ok], oit[ok])
    # The brackets the refinement started from.
    br = handle.get_brackets()
    win = prob.windows(False)
    for v in range(17):
        ref = oracle.bracket_level(on, of, orr, prob.e_lo, prob.e_top, v, oracle.band_far_exps(win), win)
        assert int(br["found"][0, v]) == ref.found
        if ref.found:
            assert (br["e_lo"][0, v], br["e_hi"][0, v]) == (ref.e_lo, ref.e_hi)
            assert (br["nodes_lo"][0, v], br["nodes_hi"][0, v]) == (ref.c_lo, ref.c_hi)


@pytest.mark.gpu
def test_float32_handle_rounds_the_float64_levels() -> None:
    """TaskConfiguratorFloat32 computes in FP64 like the float64 one and
    returns the levels rounded to float32 (the reference's FP template
    parameter sets its buffer type)."""
    m = _gpu()
    interface = m.EpseonComputeContext.create().get_device_interface(0)
    out = {}
    for precision in ("float32", "float64"):
        handle = interface.submit_task(_configure(m, interface.get_task_configurator(precision)))
        handle.wait()
        out[precision] = handle.get_levels()
    assert out["float32"].dtype == np.float32
    np.testing.assert_array_equal(out["float32"], out["float64"].astype(np.float32))


@pytest.mark.gpu
def test_drop_in_tabulated_double_wells_match_oracle() -> None:
    """Two ragged tabulated double wells through set_tabulated_potential
    with per-potential reduced masses and trial-energy counts: the handle's
    node counts and levels equal the oracle's bit for bit."""
    import oracle
    from epseon_backend_amd.workloads import double_wells

    m = _gpu()
    specs = [double_wells(64, n_points=n, n_energies=k)[i] for i, n, k in ((0, 20000, 2048), (1, 12000, 1536))]
    tables = [s.table() for s in specs]
    levels = 20
    interface = m.EpseonComputeContext.create().get_device_interface(0)
    c = (
        interface.get_task_configurator("float64")
        .set_hardware_config(potential_buffer_size=20000, group_size=2, allocation_block_size=16 << 20)
        .set_tabulated_potential([m.TabulatedPotentialConfig(values=t.tolist(), min_r=s.r_min, max_r=s.r_max,
                                                             asymptote=s.de) for s, t in zip(specs, tables)])
        .set_vibwa_algorithm(mass_atom_0=1.0, mass_atom_1=1.0, integration_step=0.1,
                             min_distance_to_asymptote=0.0, min_level=0, max_level=levels - 1)
        .set_reduced_masses([s.mu for s in specs])
        .set_trial_energy_counts([s.m for s in specs])
    )
    handle = interface.submit_task(c)
    handle.wait()
    got_levels = handle.get_levels()
    got_nodes = handle.get_node_counts()
    status = handle.get_level_status()
    # Whole-grid shots (few trial energies), segmented refinement in
    # 256-lane blocks (40 levels in the batch).
    for p, (s, t) in enumerate(zip(specs, tables)):
        prob = oracle.Problem(v=t, r_min=s.r_min, r_max=s.r_max, mu=s.mu, e_top=s.de, m=s.m)
        on, of, orr, olv, oit = oracle.solve(prob, 0, levels, 1e-10, 200, segmented=True, lanes=256, windows=False)
        np.testing.assert_array_equal(got_nodes[p], on)
        ok = ~np.isnan(olv)
        assert ok.sum() >= 10
        np.testing.assert_array_equal(status[p][ok], 0)
        np.testing.assert_array_equal(got_levels[p][ok].view(np.int64), olv[ok].view(np.int64))


def test_stub_declares_exactly_the_bound_members() -> None:
    """The .pyi stub names every public class and method the module binds,
    and nothing it does not bind."""
    import ast

    m = _gpu()
    stub = Path(m.__file__).with_name("_libepseon_gpu.pyi")
    tree = ast.parse(stub.read_text())
    declared: dict[str, set[str]] = {}
    bases: dict[str, list[str]] = {}
    for node in tree.body:
        if isinstance(node, ast.ClassDef):
            names = set()
            for item in node.body:
                if isinstance(item, ast.FunctionDef) and not item.name.startswith("__"):
                    names.add(item.name)
                elif isinstance(item, ast.AnnAssign) and isinstance(item.target, ast.Name):
                    names.add(item.target.id)
            declared[node.name] = names
            bases[node.name] = [b.id for b in node.bases if isinstance(b, ast.Name)]

    def all_names(cls: str) -> set[str]:
        out = set(declared.get(cls, set()))
        for b in bases.get(cls, []):
            out |= all_names(b)
        return out

    public = {k for k in declared if not k.startswith("_")}
    bound = {k for k in dir(m) if not k.startswith("_") and isinstance(getattr(m, k), type)}
    assert public == bound, (public ^ bound)
    for cls in public:
        have = {k for k in dir(getattr(m, cls)) if not k.startswith("_")}
        assert all_names(cls) == have, (cls, all_names(cls) ^ have)
