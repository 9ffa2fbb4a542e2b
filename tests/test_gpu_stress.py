"""Stress workload on the GPU at full size: double-well potentials (Morse plus
a Gaussian barrier) on 1,000,000-point grids with 262,144 trial energies.

The oracle's full scan at this size takes minutes on the CPU, so parity is
checked where it is exact and cheap: sampled trial energies (node count,
far-end value and rescale count bit-exact from single oracle shots), sampled
levels (the oracle refines from the GPU's own scan arrays and must land on the
same bits in the same number of iterations), eigenfunctions of sampled
levels, and the adjoint of one potential.  Properties of the whole output
(monotone counts, one level per node count below the window top, increasing
levels) cover the rest.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from epseon_backend_amd.workloads import double_wells
from tests.helpers import problem_from_spec

pytestmark = pytest.mark.gpu


def _log(msg: str) -> None:
    sys.stdout.write(msg + "\n")
    sys.stdout.flush()


@pytest.fixture(scope="module")
def capi():
    from epseon_backend_amd import capi as c

    if c.device_count() < 1:
        pytest.fail("no HIP device visible")
    return c


def test_double_wells_full_size(capi):
    specs = double_wells(64)[:2]
    _log("case: 2 double-well potentials, 1,000,000 points, 262,144 trial energies, all levels")
    tables = [s.table() for s in specs]
    counts = [s.bound_levels() + 2 for s in specs]
    rng = np.random.default_rng(17)
    with capi.Context(0) as ctx:
        ctx.load([capi.Potential(values=t, r_min=s.r_min, r_max=s.r_max, reduced_mass=s.mu, e_top=s.de,
                                 n_energies=s.m, level_min=0, level_count=c)
                  for s, t, c in zip(specs, tables, counts)])
        ctx.solve(1e-10, 200)
        nodes, far, resc = ctx.scan_results()
        lv, st, it = ctx.levels()
        stats = ctx.stats()
        _log(f"  solve {stats.solve_ms:.1f} ms, refine mode {stats.refine_mode} lanes {stats.refine_lanes}")
        psi0 = ctx.wavefunctions(0)
        g = rng.normal(size=lv.shape)
        gv = ctx.level_vjp(g)
        stats = ctx.stats()
        _log(f"  normalize {stats.normalize_ms:.1f} ms, vjp {stats.vjp_ms:.1f} ms")
    seg = stats.refine_mode == capi.REFINE_SEGMENTED
    eo = lo = 0
    for p, s in enumerate(specs):
        prob = problem_from_spec(s)
        assert prob.usable
        nd = nodes[eo : eo + s.m]
        fr = far[eo : eo + s.m]
        rs = resc[eo : eo + s.m]
        lev = lv[lo : lo + counts[p]]
        found = lev[st[lo : lo + counts[p]] == 0]
        assert np.all(np.diff(nd) >= 0)
        assert np.all(np.diff(found) > 0)
        assert len(found) == min(counts[p], int(nd[-1]))
        # Sampled trial energies: single oracle shots in the scan group's
        # band window, bit-exact.
        windows = bool(stats.scan_windows)
        for j in rng.choice(s.m, 24, replace=False):
            c, w, k = prob.scan_shot(int(j), windows)
            assert (c, k) == (int(nd[j]), int(rs[j])), (p, j)
            assert np.float64(w).view(np.int64) == fr[j].view(np.int64), (p, j)
        # Sampled levels: the oracle refines from the GPU's scan arrays.
        for v in sorted(rng.choice(len(found), 4, replace=False)):
            ref, its = prob.refine(nd, fr, rs, int(v), 1e-10, 200, segmented=seg, lanes=stats.refine_lanes,
                                   on=windows)
            assert np.float64(ref).view(np.int64) == lev[v].view(np.int64), (p, v)
            assert its == int(it[lo + v]), (p, v)
        eo += s.m
        lo += counts[p]
    # Eigenfunctions of sampled levels of potential 0 and its adjoint.
    prob = problem_from_spec(specs[0])
    for v in (0, counts[0] // 3, len(lv[: counts[0]][st[: counts[0]] == 0]) - 1):
        ref = oracle.eigenfunction(prob, lv[v])
        np.testing.assert_allclose(psi0[v], ref, rtol=0, atol=1e-12 * np.abs(ref).max(), err_msg=f"level {v}")
    lev0 = np.where(st[: counts[0]] == 0, lv[: counts[0]], np.nan)
    ref = oracle.level_vjp(prob, lev0, g[: counts[0]])
    np.testing.assert_allclose(gv[: specs[0].n], ref, rtol=0, atol=1e-11 * np.abs(ref).max())
