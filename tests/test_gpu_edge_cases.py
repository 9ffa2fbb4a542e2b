"""GPU edge cases against the oracle: tiny trial-energy counts (including one
and partial scan groups), short grids, empty and offset level ranges, levels
above the window, unresolved potentials in a batch, an empty batch, and a
context reloaded with a different batch -- with and without shot windows."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from epseon_backend_amd.workloads import MorseSpec
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


def _pot(capi, spec: MorseSpec, level_min: int, level_count: int):
    return capi.Potential(values=spec.table(), r_min=spec.r_min, r_max=spec.r_max, reduced_mass=spec.mu,
                          e_top=spec.de, n_energies=spec.m, level_min=level_min, level_count=level_count)


def _set(ctx, capi, seg: bool, windows: int) -> None:
    t = ctx.tuning()
    t.refine_mode = capi.REFINE_SEGMENTED if seg else capi.REFINE_LANE
    t.windows = windows
    ctx.set_tuning(t)


# (spec, level_min, level_count): M of 1, 5, 63, 64, 65 and 130 trial energies;
# grids of 40 to 3000 points; an empty level range; levels from 7; levels all
# above the window top (status UNBOUND); an unresolved potential (the window
# top far above the wall threshold on a coarse grid).
CASES = [
    (MorseSpec(1.0, 20000.0, 2.0, 1.0, 0.45, 4.0, 3000, 1), 0, 3),
    (MorseSpec(5.0, 15000.0, 1.5, 1.5, 0.95, 6.0, 2500, 5), 0, 6),
    (MorseSpec(20.0, 30000.0, 2.2, 2.0, 1.45, 5.0, 1800, 63), 2, 12),
    (MorseSpec(2.0, 8000.0, 1.2, 1.2, 0.65, 7.0, 900, 64), 0, 0),
    (MorseSpec(40.0, 25000.0, 1.8, 2.5, 1.95, 6.0, 2200, 65), 7, 9),
    (MorseSpec(0.6, 38000.0, 1.9, 0.74, 0.2, 5.0, 400, 130), 0, 30),
    (MorseSpec(3.0, 10000.0, 1.6, 1.0, 0.45, 5.0, 1200, 200), 400, 4),
    (MorseSpec(10.0, 20000.0, 2.0, 1.5, 1.0, 4.0, 40, 100), 0, 5),
    (MorseSpec(50.0, 40000.0, 2.5, 2.0, 0.2, 9.0, 60, 64), 0, 3),
]


@pytest.mark.parametrize("mode,windows", [("lane", 0), ("segmented", 0), ("segmented", 1)])
def test_edge_case_batch_parity(capi, mode, windows):
    seg = mode == "segmented"
    _log(f"case: {len(CASES)} edge-case potentials, {mode} refine, windows {windows}")
    with capi.Context(0) as ctx:
        _set(ctx, capi, seg, windows)
        ctx.load([_pot(capi, s, lmin, lc) for s, lmin, lc in CASES])
        ctx.solve(1e-10, 200)
        nodes, far, resc = ctx.scan_results()
        lv, st, it = ctx.levels()
        stats = ctx.stats()
        flags = [ctx.potential_info(i).flags for i in range(len(CASES))]
    assert stats.scan_windows == (windows if seg else 0)
    lanes = stats.refine_lanes
    eo = lo = 0
    n_resolved = n_unresolved = 0
    for i, (s, lmin, lc) in enumerate(CASES):
        prob = problem_from_spec(s)
        got_lv, got_st, got_it = lv[lo : lo + lc], st[lo : lo + lc], it[lo : lo + lc]
        if not prob.usable:
            assert flags[i] == 1
            assert np.all(got_st == 3) and np.all(np.isnan(got_lv))
            n_unresolved += 1
        else:
            assert flags[i] == 0
            n_resolved += 1
            on, of, orr, olv, oit = oracle.solve(prob, lmin, lc, 1e-10, 200, segmented=seg, lanes=lanes,
                                                 windows=bool(windows) and seg)
            np.testing.assert_array_equal(nodes[eo : eo + s.m], on)
            np.testing.assert_array_equal(resc[eo : eo + s.m], orr)
            np.testing.assert_array_equal(far[eo : eo + s.m].view(np.int64), of.view(np.int64))
            ok = ~np.isnan(olv)
            np.testing.assert_array_equal(np.isnan(got_lv), ~ok)
            np.testing.assert_array_equal(got_st[ok], 0)
            np.testing.assert_array_equal(got_st[~ok], 1)
            np.testing.assert_array_equal(got_lv[ok].view(np.int64), olv[ok].view(np.int64))
            np.testing.assert_array_equal(got_it[ok], oit[ok])
        eo += s.m
        lo += lc
    assert n_resolved >= 7 and n_unresolved >= 1
    # Levels from 400 of a potential with far fewer bound levels: all UNBOUND.
    off = sum(c[2] for c in CASES[:6])
    assert np.all(st[off : off + 4] == 1)


def test_empty_batch_and_reload(capi):
    _log("case: empty batch, then a reload with a different batch on the same context")
    with capi.Context(0) as ctx:
        ctx.load([])
        ctx.solve(1e-10, 200)
        lv, st, it = ctx.levels()
        assert lv.size == 0
        first = [CASES[2], CASES[4]]
        ctx.load([_pot(capi, s, lmin, lc) for s, lmin, lc in first])
        ctx.solve(1e-10, 200)
        a = ctx.levels()[0].copy()
        second = [CASES[5], CASES[0]]
        ctx.load([_pot(capi, s, lmin, lc) for s, lmin, lc in second])
        ctx.solve(1e-10, 200)
        b, stb, _ = ctx.levels()
        stats = ctx.stats()
    seg = stats.refine_mode == capi.REFINE_SEGMENTED
    for got, batch in ((a, first), (b, second)):
        lo = 0
        for s, lmin, lc in batch:
            prob = problem_from_spec(s)
            olv = oracle.solve(prob, lmin, lc, 1e-10, 200, segmented=seg, lanes=stats.refine_lanes,
                               windows=bool(stats.scan_windows))[3]
            ok = ~np.isnan(olv)
            np.testing.assert_array_equal(got[lo : lo + lc][ok].view(np.int64), olv[ok].view(np.int64))
            lo += lc
