"""Single heavy Morse potential at full size on the GPU (the I2-like workload).

mu = 63.452 amu, D_e = 12440 cm^-1, a = 1.8676 1/A, r_e = 2.666 A on a
200,000-point FP64 grid over [2, 8] A with 65,536 trial energies: the whole
scan (node counts, far-end mantissas, rescale counts of every trial energy)
and every level with its iteration count bit-exact against the oracle in the
mode and block width the library chose, and the levels well inside the box
against the analytic Morse levels of tests/golden/morse_levels.json.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle
from epseon_backend_amd.workloads import I2

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def _log(msg: str) -> None:
    sys.stdout.write(msg + "\n")
    sys.stdout.flush()


@pytest.fixture(scope="module")
def capi():
    from epseon_backend_amd import capi as c

    if c.device_count() < 1:
        pytest.fail("no HIP device visible")
    return c


def _solve(capi, levels: int, windows: int | None = None):
    table = I2.table()
    with capi.Context(0) as ctx:
        if windows is not None:
            t = ctx.tuning()
            t.windows = windows
            ctx.set_tuning(t)
        ctx.load([capi.Potential(values=table, r_min=I2.r_min, r_max=I2.r_max, reduced_mass=I2.mu, e_top=I2.de,
                                 n_energies=I2.m, level_min=0, level_count=levels)])
        ctx.solve(1e-10, 200)
        nodes, far, resc = ctx.scan_results()
        lv, st, it = ctx.levels()
        stats = ctx.stats()
    return table, nodes, far, resc, lv, st, it, stats


@pytest.mark.parametrize("windows", [None, 1])
def test_i2_full_size_bit_exact(capi, windows):
    levels = 120
    label = "automatic (whole grid)" if windows is None else "band windows forced on"
    _log(f"case: I2 200000 points x 65536 trial energies, {levels} levels requested, shots: {label}")
    table, nodes, far, resc, lv, st, it, stats = _solve(capi, levels, windows)
    seg = stats.refine_mode == capi.REFINE_SEGMENTED
    _log(f"  refine mode {stats.refine_mode}, lanes {stats.refine_lanes}, windows {stats.scan_windows}, "
         f"solve {stats.solve_ms:.3f} ms, converged {stats.levels_converged}")
    if windows is None:
        # One scan wavefront per SIMD: whole-grid shots, 256-lane segmented blocks.
        assert stats.scan_windows == 0 and seg and stats.refine_lanes == 256
    prob = oracle.Problem(v=table, r_min=I2.r_min, r_max=I2.r_max, mu=I2.mu, e_top=I2.de, m=I2.m)
    on, of, orr, olv, oit = oracle.solve(prob, 0, levels, 1e-10, 200, segmented=seg, lanes=stats.refine_lanes,
                                         windows=bool(stats.scan_windows))
    np.testing.assert_array_equal(nodes, on)
    np.testing.assert_array_equal(resc, orr)
    np.testing.assert_array_equal(far.view(np.int64), of.view(np.int64))
    found = ~np.isnan(olv)
    np.testing.assert_array_equal(np.isnan(lv), ~found)
    np.testing.assert_array_equal(lv[found].view(np.int64), olv[found].view(np.int64))
    np.testing.assert_array_equal(it, oit)
    # The top trial energy (12439.81 cm^-1) has 115 nodes below it, and every
    # one of those levels converged; the 116th level lies above it.
    assert int(np.sum(st == capi.LEVEL_OK)) == int(nodes[-1]) == 115
    assert np.all(st[115:] == capi.LEVEL_UNBOUND)
    golden = np.array([float(x) for x in json.loads((GOLDEN / "morse_levels.json").read_text())["i2"]])
    # Refinement tolerance 1e-10 relative plus the ~1e-11 discretisation error.
    np.testing.assert_allclose(lv[:100], golden[:100], rtol=1e-10)
    assert np.all(np.diff(lv[:115]) > 0)
