"""The drop-in task API on the GPU beyond the reference's own contract tests:
cancellation and the running flag (reference TaskHandle::cancel/isRunning,
cpp/gpu/include/epseon/gpu/task_handle.hpp:117-144), potentials from
set_double_well_potential and from a PotentialFileLoader
(potential_source.hpp:43-100) bit-exact against the oracle or the tabulated
path, and group_size (hardware_config.hpp:11-76) smaller than the batch
giving the same bits as one group.
"""
from __future__ import annotations

import math
import re
import sys
import time

import numpy as np
import pytest

import oracle
from epseon_backend_amd.workloads import H2, batched, double_wells

pytestmark = pytest.mark.gpu


def _log(msg: str) -> None:
    sys.stdout.write(msg + "\n")
    sys.stdout.flush()


@pytest.fixture(scope="module")
def m():
    from epseon_backend_amd.device.gpu import _libepseon_gpu as mod

    if len(mod.EpseonComputeContext.create().get_physical_device_info()) < 1:
        pytest.fail("no HIP device visible")
    return mod


@pytest.fixture(scope="module")
def interface(m):
    return m.EpseonComputeContext.create().get_device_interface(0)


def _double_well_configs(m, specs):
    return [m.DoubleWellPotentialConfig(dissociation_energy=s.de, equilibrium_bond_distance=s.re, well_width=s.a,
                                        barrier_height=s.barrier, barrier_position=s.r_b, barrier_width=s.width,
                                        min_r=s.r_min, max_r=s.r_max, point_count=s.n) for s in specs]


def _heavy_task(m, interface, group_size):
    specs = double_wells(64, n_points=1_000_000, n_energies=4096)[:24]
    return (
        interface.get_task_configurator("float64")
        .set_hardware_config(potential_buffer_size=1_000_000, group_size=group_size, allocation_block_size=1 << 26)
        .set_double_well_potential(_double_well_configs(m, specs))
        .set_vibwa_algorithm(mass_atom_0=2 * specs[0].mu, mass_atom_1=2 * specs[0].mu, integration_step=0.1,
                             min_distance_to_asymptote=0.0, min_level=0, max_level=9)
    )


def test_cancel_before_the_device_phases(m, interface):
    """A task cancelled right after submit stops at its next phase boundary
    (it is still tabulating 24 million grid points): wait() returns, the
    status says so, and the results raise."""
    _log("case: cancel right after submit (24 double wells x 1,000,000 points)")
    handle = interface.submit_task(_heavy_task(m, interface, 0))
    assert handle.is_running()
    assert handle.cancel() is True
    handle.wait()
    assert handle.is_done() and not handle.is_running()
    assert handle.get_status_message() == "cancelled"
    with pytest.raises(RuntimeError, match="cancelled"):
        handle.get_levels()
    assert handle.cancel() is False


def test_cancel_between_groups(m, interface):
    """With group_size 1 the worker loads and solves one potential at a
    time; a cancel issued once the third group is under way stops it before
    the batch ends."""
    _log("case: cancel during group 3 of 24 (group_size 1)")
    handle = interface.submit_task(_heavy_task(m, interface, 1))
    t0 = time.perf_counter()
    group = -1
    while time.perf_counter() - t0 < 60.0 and not handle.is_done():
        g = re.search(r"potentials (\d+)\.\.", handle.get_status_message())
        if g:
            group = int(g.group(1))
            if group >= 2:
                break
        time.sleep(0.0005)
    assert group >= 2 and not handle.is_done(), handle.get_status_message()
    assert handle.is_running()
    assert handle.cancel() is True
    handle.wait()
    assert handle.get_status_message() == "cancelled"
    with pytest.raises(RuntimeError, match="cancelled"):
        handle.get_level_status()


def test_double_wells_through_the_drop_in_match_oracle(m, interface):
    """set_double_well_potential tabulates V = D_e (1 - e^{-a(r - r_e)})^2 +
    B e^{-((r - r_b)/w)^2} with the C library's exp; the same table built
    here with math.exp gives the oracle identical bytes."""
    specs = double_wells(64, n_points=20000, n_energies=2048)[:2]
    levels = 16
    _log("case: 2 double wells x 20000 points through set_double_well_potential")
    c = (
        interface.get_task_configurator("float64")
        .set_hardware_config(potential_buffer_size=20000, group_size=2, allocation_block_size=1 << 24)
        .set_double_well_potential(_double_well_configs(m, specs))
        .set_vibwa_algorithm(mass_atom_0=1.0, mass_atom_1=1.0, integration_step=0.1,
                             min_distance_to_asymptote=0.0, min_level=0, max_level=levels - 1)
        .set_reduced_masses([s.mu for s in specs])
        .set_trial_energies(2048)
    )
    handle = interface.submit_task(c)
    handle.wait()
    lv = handle.get_levels()
    nodes = handle.get_node_counts()
    far = handle.get_far_end_values()
    its = handle.get_iterations()
    for p, s in enumerate(specs):
        h = (s.r_max - s.r_min) / (s.n - 1)
        v = np.empty(s.n)
        for i in range(s.n):
            r = s.r_min + i * h
            x = 1.0 - math.exp(-s.a * (r - s.re))
            z = (r - s.r_b) / s.width
            v[i] = s.de * (x * x) + s.barrier * math.exp(-(z * z))
        prob = oracle.Problem(v=v, r_min=s.r_min, r_max=s.r_max, mu=s.mu, e_top=s.de, m=2048)
        on, of, orr, olv, oit = oracle.solve(prob, 0, levels, 1e-10, 200, segmented=True, lanes=256, windows=False)
        np.testing.assert_array_equal(nodes[p], on)
        np.testing.assert_array_equal(far[p].view(np.int64), of.view(np.int64))
        ok = ~np.isnan(olv)
        assert ok.sum() >= 8
        np.testing.assert_array_equal(lv[p][ok].view(np.int64), olv[ok].view(np.int64))
        np.testing.assert_array_equal(its[p][ok], oit[ok])


def test_group_size_smaller_than_the_batch_is_bit_identical(m, interface):
    """group_size 1 (one load and solve per potential) against one group:
    levels, iterations, status, node counts, far ends, brackets and
    eigenfunctions identical; the handle's arrays are read-only views."""
    specs = batched(4096)[:5]
    levels = 20
    cfgs = [m.MorsePotentialConfig(dissociation_energy=s.de, equilibrium_bond_distance=s.re, well_width=s.a,
                                   min_r=s.r_min, max_r=s.r_max, point_count=s.n) for s in specs]
    _log("case: 5 ragged batch potentials, group_size 1 vs 5, with eigenfunctions")
    out = []
    for group in (1, 5):
        c = (
            interface.get_task_configurator("float64")
            .set_hardware_config(potential_buffer_size=64000, group_size=group, allocation_block_size=1 << 24)
            .set_morse_potential_batch(cfgs)
            .set_vibwa_algorithm(mass_atom_0=1.0, mass_atom_1=1.0, integration_step=0.1,
                                 min_distance_to_asymptote=0.0, min_level=0, max_level=levels - 1)
            .set_reduced_masses([s.mu for s in specs])
            .set_trial_energy_counts([s.m for s in specs])
            .set_eigenfunction_output(True)
        )
        handle = interface.submit_task(c)
        handle.wait()
        psi = handle.get_wavefunctions()
        assert not psi[0].flags.writeable
        np.testing.assert_array_equal(handle.get_wavefunctions(potential=3), psi[3])
        br = handle.get_brackets()
        out.append((handle.get_levels().view(np.int64), handle.get_iterations(), handle.get_level_status(),
                    [x.copy() for x in handle.get_node_counts()],
                    [x.view(np.int64).copy() for x in handle.get_far_end_values()],
                    [x.copy() for x in handle.get_rescale_counts()],
                    br["e_lo"].view(np.int64), br["e_hi"].view(np.int64), br["nodes_lo"], br["nodes_hi"],
                    br["found"], [x.view(np.int64).copy() for x in psi]))
    for a, b in zip(out[0], out[1]):
        if isinstance(a, list):
            for x, y in zip(a, b):
                np.testing.assert_array_equal(x, y)
        else:
            np.testing.assert_array_equal(a, b)
    assert int(np.sum(out[0][2] == 0)) > 40


def test_file_loader_equals_tabulated(m, interface, tmp_path):
    """Potentials read from a text file and a .npy file give the same bits
    as set_tabulated_potential on the same values."""
    n = H2.n
    h = (H2.r_max - H2.r_min) / (n - 1)
    r = H2.r_min + np.arange(n) * h
    v = H2.table()
    txt = tmp_path / "h2.txt"
    txt.write_text(f"# asymptote {float(H2.de)!r}\n" + "".join(f"{float(a)!r} {float(b)!r}\n" for a, b in zip(r, v)))
    npy = tmp_path / "h2.npy"
    np.save(npy, np.column_stack([r, v]))
    _log("case: H2 table from a text file and a .npy file vs set_tabulated_potential")

    def run(setter):
        c = interface.get_task_configurator("float64").set_hardware_config(
            potential_buffer_size=n, group_size=2, allocation_block_size=1 << 24)
        setter(c)
        c.set_vibwa_algorithm(mass_atom_0=2 * H2.mu, mass_atom_1=2 * H2.mu, integration_step=0.1,
                              min_distance_to_asymptote=0.0, min_level=0, max_level=15)
        c.set_trial_energies(2048)
        handle = interface.submit_task(c)
        handle.wait()
        return handle.get_levels(), handle.get_iterations()

    from_files = run(lambda c: c.set_potential_file_loader(m.PotentialFileLoader([str(txt), str(npy)])))
    # The text file's asymptote line and the .npy's last value set the top.
    tab = run(lambda c: c.set_tabulated_potential([
        m.TabulatedPotentialConfig(values=v.tolist(), min_r=float(r[0]), max_r=float(r[-1]), asymptote=H2.de),
        m.TabulatedPotentialConfig(values=v.tolist(), min_r=float(r[0]), max_r=float(r[-1]),
                                   asymptote=float(v[-1]))]))
    np.testing.assert_array_equal(from_files[0].view(np.int64), tab[0].view(np.int64))
    np.testing.assert_array_equal(from_files[1], tab[1])
    assert np.sum(np.isfinite(from_files[0][0])) >= 14
