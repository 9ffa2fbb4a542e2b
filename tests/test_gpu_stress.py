"""Stress workload on the GPU at full size: all 64 double-well potentials
(Morse plus a Gaussian barrier) on 1,000,000-point grids with 262,144 trial
energies, every bound level, every eigenfunction written to device memory
and the adjoint of the whole batch.

The oracle's full scan at this size takes minutes on the CPU, so parity is
checked where it is exact and cheap: sampled trial energies of every
potential (node count, far-end value and rescale count bit-exact from single
oracle shots), sampled levels of every potential (the oracle refines from the
GPU's own scan arrays and must land on the same bits in the same number of
iterations), eigenfunctions of sampled levels and the adjoint of sampled
potentials.  Properties of the whole output (monotone counts, one level per
node count below the window top, increasing levels, unit-norm eigenfunctions)
cover the rest.
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
    import torch

    specs = double_wells(64)
    _log("case: all 64 double-well potentials, 1,000,000 points, 262,144 trial energies, all levels")
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
        _log(f"  solve {stats.solve_ms:.1f} ms, refine mode {stats.refine_mode} lanes {stats.refine_lanes}, "
             f"windows {stats.scan_windows}, converged {stats.levels_converged}")
        # Every eigenfunction of the batch in device memory (one launch), and
        # the adjoint of the whole batch with eigenfunctions recomputed.
        psi = torch.empty(ctx.eigenfunction_size(), dtype=torch.float64, device="cuda")
        ctx.all_wavefunctions_device(psi.data_ptr())
        g = rng.normal(size=lv.shape)
        gd = torch.from_numpy(g).cuda()
        gv_dev = torch.empty(ctx.v_span, dtype=torch.float64, device="cuda")
        ctx.level_vjp_device(gd.data_ptr(), gv_dev.data_ptr())
        stats = ctx.stats()
        _log(f"  normalize {stats.normalize_ms:.1f} ms, vjp {stats.vjp_ms:.1f} ms")
    gv = gv_dev.cpu().numpy()
    seg = stats.refine_mode == capi.REFINE_SEGMENTED
    windows = bool(stats.scan_windows)
    eo = lo = po = 0
    psi_base = []
    for p, s in enumerate(specs):
        prob = problem_from_spec(s)
        prob.v = tables[p]
        assert prob.usable
        nd = nodes[eo : eo + s.m]
        fr = far[eo : eo + s.m]
        rs = resc[eo : eo + s.m]
        lev = lv[lo : lo + counts[p]]
        ok = st[lo : lo + counts[p]] == 0
        found = lev[ok]
        assert np.all(np.diff(nd) >= 0), p
        assert np.all(np.diff(found) > 0), p
        assert len(found) == min(counts[p], int(nd[-1])), p
        assert np.all(ok[: len(found)]), p
        # Sampled trial energies: single oracle shots in the scan group's
        # band window, bit-exact.
        for j in rng.choice(s.m, 6, replace=False):
            c, w, k = prob.scan_shot(int(j), windows)
            assert (c, k) == (int(nd[j]), int(rs[j])), (p, j)
            assert np.float64(w).view(np.int64) == fr[j].view(np.int64), (p, j)
        # Sampled levels: the oracle refines from the GPU's scan arrays.
        for v in sorted(rng.choice(len(found), 2, replace=False)):
            ref, its = prob.refine(nd, fr, rs, int(v), 1e-10, 200, segmented=seg, lanes=stats.refine_lanes,
                                   on=windows)
            assert np.float64(ref).view(np.int64) == lev[v].view(np.int64), (p, v)
            assert its == int(it[lo + v]), (p, v)
        psi_base.append(po)
        eo += s.m
        lo += counts[p]
        po += counts[p] * s.n
    assert po == psi.numel()
    # Eigenfunctions of sampled levels of sampled potentials against the
    # oracle at the GPU's levels; unit norm on the grid.
    lo_of = np.concatenate([[0], np.cumsum(counts)])
    for p in sorted(rng.choice(len(specs), 4, replace=False)):
        s = specs[p]
        prob = problem_from_spec(s)
        prob.v = tables[p]
        nf = int(np.sum(st[lo_of[p] : lo_of[p + 1]] == 0))
        for v in (0, nf // 2, nf - 1):
            row = psi[psi_base[p] + v * s.n : psi_base[p] + (v + 1) * s.n].cpu().numpy()
            ref = oracle.eigenfunction(prob, lv[lo_of[p] + v])
            np.testing.assert_allclose(row, ref, rtol=0, atol=1e-12 * np.abs(ref).max(), err_msg=f"{p}/{v}")
            assert abs(float(np.sum(row * row) * prob.h) - 1.0) < 1e-12
    # The adjoint of sampled potentials against the oracle's.
    off = np.concatenate([[0], np.cumsum([s.n for s in specs])])
    for p in sorted(rng.choice(len(specs), 2, replace=False)):
        s = specs[p]
        prob = problem_from_spec(s)
        prob.v = tables[p]
        levp = np.where(st[lo_of[p] : lo_of[p + 1]] == 0, lv[lo_of[p] : lo_of[p + 1]], np.nan)
        ref = oracle.level_vjp(prob, levp, g[lo_of[p] : lo_of[p + 1]])
        np.testing.assert_allclose(gv[off[p] : off[p + 1]], ref, rtol=0, atol=1e-11 * np.abs(ref).max(),
                                   err_msg=f"potential {p}")
