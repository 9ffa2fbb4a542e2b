"""PotentialFileLoader: the reference's file-based potential source
(cpp/gpu/include/epseon/gpu/task_configurator/potential_source.hpp:43-100),
whose data getter is a stub there, with a file format of this module's own:
text "r V" columns (optional "# asymptote <x>") or (N, 2) float64 .npy.

The reference pins only that a default-constructed loader has no data and
that loaders compare by file names (cpp/gpu/test/task_configurator/
test_potential_source.cpp:23-56); the values read here must equal the file's
bit for bit, and malformed files must raise.
"""
from __future__ import annotations

import numpy as np
import pytest

from epseon_backend_amd.workloads import H2

gpu_mod = pytest.importorskip("epseon_backend_amd.device.gpu._libepseon_gpu")


def _write_text(path, r, v, asymptote=None, comma=False):
    lines = ["# H2-like Morse potential: r [A], V [cm^-1]"]
    if asymptote is not None:
        lines.append(f"# asymptote {float(asymptote)!r}")
    sep = ", " if comma else " "
    # repr of a Python float round-trips exactly.
    lines += [f"{float(x)!r}{sep}{float(y)!r}" for x, y in zip(r, v)]
    path.write_text("\n".join(lines) + "\n")


def _grid(spec, n=None):
    n = n or spec.n
    h = (spec.r_max - spec.r_min) / (n - 1)
    r = spec.r_min + np.arange(n) * h
    x = 1.0 - np.exp(-spec.a * (r - spec.re))
    return r, spec.de * x * x


def test_default_loader_has_no_data():
    loader = gpu_mod.PotentialFileLoader()
    assert loader.get_potential_data() == []
    assert loader.file_names == []


def test_loaders_compare_by_file_names():
    a = gpu_mod.PotentialFileLoader(["file1.txt", "file2.txt"])
    b = gpu_mod.PotentialFileLoader(["file1.txt", "file2.txt"])
    c = gpu_mod.PotentialFileLoader(["file1.txt"])
    assert a.equals(b) and a == b
    assert not a.equals(c) and not a == c
    assert a.file_names == ["file1.txt", "file2.txt"]


def test_text_and_npy_values_read_bit_for_bit(tmp_path):
    r, v = _grid(H2, 1000)
    t1 = tmp_path / "h2.txt"
    t2 = tmp_path / "h2_comma.dat"
    n1 = tmp_path / "h2.npy"
    _write_text(t1, r, v, asymptote=38297.0)
    _write_text(t2, r, v, comma=True)
    np.save(n1, np.column_stack([r, v]))
    data = gpu_mod.PotentialFileLoader([str(t1), str(t2), str(n1)]).get_potential_data()
    assert len(data) == 3
    for d in data:
        assert d.dtype == np.float64
        np.testing.assert_array_equal(d.view(np.int64), v.view(np.int64))


@pytest.mark.parametrize(
    "body, message",
    [
        ("1.0 2.0\n2.0 3.0\n3.0 4.0\n", "at least 4"),
        ("1.0 2.0\n2.0 3.0\n3.5 4.0\n4.0 5.0\n", "uniform"),
        ("1.0 2.0\n2.0 nan\n3.0 4.0\n4.0 5.0\n", "not finite"),
        ("1.0 2.0\n2.0\n3.0 4.0\n4.0 5.0\n", "line 2"),
        ("1.0 2.0 7.0\n2.0 3.0\n3.0 4.0\n4.0 5.0\n", "unexpected text"),
        ("4.0 2.0\n3.0 3.0\n2.0 4.0\n1.0 5.0\n", "increase"),
        ("# asymptote x\n1.0 2.0\n2.0 3.0\n3.0 4.0\n4.0 5.0\n", "asymptote"),
        ("# only comments\n\n", "no data"),
    ],
)
def test_malformed_text_files_raise(tmp_path, body, message):
    f = tmp_path / "bad.txt"
    f.write_text(body)
    with pytest.raises(RuntimeError, match=message):
        gpu_mod.PotentialFileLoader([str(f)]).get_potential_data()


def test_bad_npy_files_and_missing_files_raise(tmp_path):
    f32 = tmp_path / "f32.npy"
    np.save(f32, np.ones((8, 2), dtype=np.float32))
    shape = tmp_path / "shape.npy"
    np.save(shape, np.ones((8, 3)))
    fortran = tmp_path / "fortran.npy"
    np.save(fortran, np.asfortranarray(np.column_stack([np.arange(8.0), np.ones(8)])))
    for path, msg in ((f32, "float64"), (shape, r"\(N, 2\)"), (fortran, "C order")):
        with pytest.raises(RuntimeError, match=msg):
            gpu_mod.PotentialFileLoader([str(path)]).get_potential_data()
    with pytest.raises(RuntimeError, match="cannot be opened"):
        gpu_mod.PotentialFileLoader([str(tmp_path / "missing.txt")]).get_potential_data()


def test_configurator_reads_files_and_returns_itself(tmp_path):
    r, v = _grid(H2, 500)
    f = tmp_path / "h2.txt"
    _write_text(f, r, v, asymptote=38297.0)
    for precision in ("float32", "float64"):
        cfg = gpu_mod.TaskConfiguratorFloat64() if precision == "float64" else gpu_mod.TaskConfiguratorFloat32()
        same = cfg.set_potential_file_loader(gpu_mod.PotentialFileLoader([str(f)]))
        assert same is cfg
        assert not cfg.is_configured()
        cfg.set_hardware_config(potential_buffer_size=1000, group_size=1, allocation_block_size=1 << 24)
        cfg.set_vibwa_algorithm(mass_atom_0=1.00782, mass_atom_1=1.00782, integration_step=0.001,
                                min_distance_to_asymptote=0.0, min_level=0, max_level=3)
        assert cfg.is_configured()
    with pytest.raises(RuntimeError, match="cannot be opened"):
        gpu_mod.TaskConfiguratorFloat64().set_potential_file_loader(
            gpu_mod.PotentialFileLoader([str(tmp_path / "nope.txt")]))
