"""Batched workload on the GPU: 4,096 seeded Morse potentials with ragged grids.

* a subset of the batch at full size, bit-exact against the oracle, in both
  refinement modes;
* the whole batch through the default (auto) mode, checked by properties
  that do not depend on size (monotone counts, increasing levels, one level
  per node count below the top trial energy, analytic Morse agreement where
  the grid resolves the well) plus a random sample of potentials bit-exact
  against the oracle.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from epseon_backend_amd.workloads import batched
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


def _pot(capi, spec, table, levels):
    return capi.Potential(values=table, r_min=spec.r_min, r_max=spec.r_max, reduced_mass=spec.mu,
                          e_top=spec.de, n_energies=spec.m, level_min=0, level_count=levels)


def _compare(prob, nodes, far, resc, lv, it, levels, seg, lanes, windows=None):
    # Lane mode shoots the whole grid; segmented mode the band windows.
    windows = seg if windows is None else windows
    on, of, orr, olv, oit = oracle.solve(prob, 0, levels, 1e-10, 200, segmented=seg, lanes=lanes,
                                         windows=windows)
    np.testing.assert_array_equal(nodes, on)
    np.testing.assert_array_equal(resc, orr)
    np.testing.assert_array_equal(far.view(np.int64), of.view(np.int64))
    ok = ~np.isnan(olv)
    np.testing.assert_array_equal(np.isnan(lv), ~ok)
    np.testing.assert_array_equal(lv[ok].view(np.int64), olv[ok].view(np.int64))
    np.testing.assert_array_equal(it[ok], oit[ok])


@pytest.mark.parametrize("mode", ["lane", "segmented", "segmented-pairs", "segmented-quads"])
def test_batched_subset_parity(capi, mode):
    specs = batched(4096)[:40]
    _log(f"case: first 40 potentials of the batched workload at full size, {mode} refine")
    seg = mode != "lane"
    tables = [s.table() for s in specs]
    counts = [s.bound_levels() + 2 for s in specs]
    with capi.Context(0) as ctx:
        # Segmented with band windows (forced: 40 potentials are below the
        # automatic threshold), lane over the whole grid.
        t = ctx.tuning()
        t.refine_mode = capi.REFINE_SEGMENTED if seg else capi.REFINE_LANE
        t.windows = int(seg)
        # Two or four levels per 64-lane block (each table load feeding the
        # shots of every level of the group in the same window); every
        # level's arithmetic is that of a block of its own.
        lpb = {"segmented-pairs": 2, "segmented-quads": 4}.get(mode, 1)
        t.segmented_levels = lpb
        ctx.set_tuning(t)
        ctx.load([_pot(capi, s, t, c) for s, t, c in zip(specs, tables, counts)])
        ctx.solve(1e-10, 200)
        nodes, far, resc = ctx.scan_results()
        lv, st, it = ctx.levels()
        lanes = ctx.stats().refine_lanes
        assert ctx.stats().scan_windows == int(seg)
        if seg:
            assert ctx.stats().refine_levels_per_block == lpb
        flags = [ctx.potential_info(i).flags for i in range(len(specs))]
    eo = lo = 0
    for i, s in enumerate(specs):
        prob = problem_from_spec(s)
        prob.v = tables[i]
        assert flags[i] == (0 if prob.usable else 1)
        if flags[i] == 0:
            _compare(prob, nodes[eo : eo + s.m], far[eo : eo + s.m], resc[eo : eo + s.m],
                     lv[lo : lo + counts[i]], it[lo : lo + counts[i]], counts[i], seg, lanes)
        else:
            assert np.all(st[lo : lo + counts[i]] == 3)
        eo += s.m
        lo += counts[i]


def test_batched_full_scale(capi):
    specs = batched(4096)
    _log("case: full batched workload, 4096 potentials, auto refine mode")
    tables = [s.table() for s in specs]
    counts = [s.bound_levels() + 2 for s in specs]
    with capi.Context(0) as ctx:
        ctx.load([_pot(capi, s, t, c) for s, t, c in zip(specs, tables, counts)])
        ctx.solve(1e-10, 200)
        nodes, far, resc = ctx.scan_results()
        lv, st, it = ctx.levels()
        stats = ctx.stats()
        flags = np.array([ctx.potential_info(i).flags for i in range(len(specs))])
    _log(f"  refine mode {stats.refine_mode}, lanes {stats.refine_lanes}, solve {stats.solve_ms:.2f} ms, "
         f"refine {stats.refine_ms:.2f} ms, levels converged {stats.levels_converged}, "
         f"unresolved potentials {int(flags.sum())}")
    assert stats.refine_mode == capi.REFINE_SEGMENTED and stats.refine_lanes == 64 and stats.scan_windows == 1
    assert stats.refine_levels_per_block == 4
    assert np.all(st != 2), "a level hit the iteration cap"
    eo = lo = 0
    resolved_checked = 0
    starts_e, starts_l = [], []
    for i, s in enumerate(specs):
        starts_e.append(eo)
        starts_l.append(lo)
        nd = nodes[eo : eo + s.m]
        lev = lv[lo : lo + counts[i]]
        if flags[i] == 0:
            assert np.all(np.diff(nd) >= 0), f"potential {i}: node counts not monotone"
            found = lev[~np.isnan(lev)]
            assert np.all(np.diff(found) > 0), f"potential {i}: levels not increasing"
            # Sturm property: the node count at the top trial energy (D_e) is
            # the number of box levels below it, and every one of them that was
            # requested is found.  The box (Dirichlet wall at V(r_max) =
            # 0.995 D_e) shifts the top levels, so the analytic Morse count is
            # only a lower bound well below the top.
            assert len(found) == min(counts[i], int(nd[-1])), (i, len(found), int(nd[-1]))
            nb_low = int(np.sum(s.analytic(np.arange(s.bound_levels())) < 0.9 * s.de))
            assert int(nd[-1]) >= nb_low - 1, (i, int(nd[-1]), nb_low)
            # Low levels against analytic Morse where the grid resolves the well
            # and the wall at r_e - 0.6 A is opaque (WKB: 2 int kappa > 40 from
            # the wall to the inner turning point of v = 2).
            h = (s.r_max - s.r_min) / (s.n - 1)
            kh = np.sqrt(s.de * s.mu / 16.857629191640175) * h
            ana = s.analytic(np.arange(3))
            vt = tables[i]
            left = vt[: int(np.argmin(vt)) + 1]
            kappa = np.sqrt(np.clip(left - ana[2], 0.0, None) * s.mu / 16.857629191640175)
            if kh < 0.2 and 2.0 * kappa.sum() * h > 40.0:
                np.testing.assert_allclose(found[:3], ana, rtol=max(1e-8, kh**4 / 10))
                resolved_checked += 1
        eo += s.m
        lo += counts[i]
    _log(f"  analytic Morse check on {resolved_checked} potentials")
    assert resolved_checked > 1500
    rng = np.random.default_rng(1)
    for i in rng.choice(len(specs), 6, replace=False):
        if flags[i]:
            continue
        s = specs[i]
        prob = problem_from_spec(s)
        prob.v = tables[i]
        e0, l0 = starts_e[i], starts_l[i]
        _compare(prob, nodes[e0 : e0 + s.m], far[e0 : e0 + s.m], resc[e0 : e0 + s.m],
                 lv[l0 : l0 + counts[i]], it[l0 : l0 + counts[i]], counts[i], True, 64)


def test_result_preserving_tunings_are_bit_identical(capi):
    """Scan workgroup size, XCD binning, one or two levels per lane and the
    two-stream pipelined solve only change the schedule: every output is
    bit-identical."""
    specs = batched(4096)[:600]
    _log("case: first 600 batched potentials, lane mode, tunings varied")
    counts = [s.bound_levels() + 2 for s in specs]
    outs = []
    with capi.Context(0) as ctx:
        ctx.load([_pot(capi, s, s.table(), c) for s, c in zip(specs, counts)])
        for wpb, lpl, split, xcd in ((1, 1, 0.0, 1), (2, 2, 0.0, 0), (4, 1, 0.3, 1), (1, 2, 0.7, 0),
                                     (2, 1, 0.0, 1)):
            t = capi.Tuning(refine_mode=capi.REFINE_LANE, levels_per_lane=lpl, scan_waves_per_block=wpb,
                            segmented_lanes=0, xcd_scan=xcd, windows=0, solve_split=split)
            ctx.set_tuning(t)
            ctx.solve(1e-10, 200)
            nodes, far, resc = ctx.scan_results()
            lv, st, it = ctx.levels()
            stats = ctx.stats()
            _log(f"  wpb {wpb} levels/lane {lpl} split {split} xcd {xcd}: solve {stats.solve_ms:.2f} ms")
            assert (stats.solve_split > 0) == (split > 0)
            outs.append((nodes, far.view(np.int64), resc, lv.view(np.int64), st, it))
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            np.testing.assert_array_equal(a, b)


def test_result_preserving_tunings_windowed(capi):
    """The same schedule knobs with windowed shots and segmented refinement,
    one, two or four levels per block: bit-identical outputs, equal to the
    oracle's windowed solve."""
    specs = batched(4096)[:300]
    _log("case: first 300 batched potentials, segmented refine with band windows, tunings varied")
    counts = [s.bound_levels() + 2 for s in specs]
    tables = [s.table() for s in specs]
    outs = []
    with capi.Context(0) as ctx:
        ctx.load([_pot(capi, s, t, c) for s, t, c in zip(specs, tables, counts)])
        for wpb, split, xcd, lpb in ((1, 0.0, 1, 1), (2, 0.3, 0, 2), (4, 0.0, 0, 1), (1, 0.0, 1, 2),
                                     (2, 0.0, 0, 2), (1, 0.0, 1, 4), (2, 0.3, 0, 4)):
            t = capi.Tuning(refine_mode=capi.REFINE_SEGMENTED, levels_per_lane=1, scan_waves_per_block=wpb,
                            segmented_lanes=64, xcd_scan=xcd, windows=1, segmented_levels=lpb,
                            solve_split=split)
            ctx.set_tuning(t)
            ctx.solve(1e-10, 200)
            nodes, far, resc = ctx.scan_results()
            lv, st, it = ctx.levels()
            stats = ctx.stats()
            assert stats.refine_levels_per_block == lpb
            _log(f"  wpb {wpb} split {split} xcd {xcd} levels/block {lpb}: solve {stats.solve_ms:.2f} ms")
            outs.append((nodes, far.view(np.int64), resc, lv.view(np.int64), st, it))
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            np.testing.assert_array_equal(a, b)
    nodes, far, resc, lvi, st, it = outs[0]
    lv = lvi.view(np.float64)
    eo = lo = 0
    for i, s in enumerate(specs[:40]):
        prob = problem_from_spec(s)
        prob.v = tables[i]
        if prob.usable:
            _compare(prob, nodes[eo : eo + s.m], far.view(np.float64)[eo : eo + s.m], resc[eo : eo + s.m],
                     lv[lo : lo + counts[i]], it[lo : lo + counts[i]], counts[i], True, 64)
        eo += s.m
        lo += counts[i]


def test_autotune_then_parity(capi):
    specs = batched(4096)[:40]
    _log("case: autotune on the first 40 batched potentials, then oracle parity in the chosen mode")
    counts = [s.bound_levels() + 2 for s in specs]
    tables = [s.table() for s in specs]
    with capi.Context(0) as ctx:
        ctx.load([_pot(capi, s, t, c) for s, t, c in zip(specs, tables, counts)])
        best = ctx.autotune(1e-10, 200)
        _log(f"  chosen: {best.as_dict()}")
        assert ctx.tuning().as_dict() == best.as_dict()
        ctx.solve(1e-10, 200)
        nodes, far, resc = ctx.scan_results()
        lv, st, it = ctx.levels()
        stats = ctx.stats()
    seg = stats.refine_mode == capi.REFINE_SEGMENTED
    eo = lo = 0
    for i, s in enumerate(specs):
        prob = problem_from_spec(s)
        prob.v = tables[i]
        if prob.usable:
            _compare(prob, nodes[eo : eo + s.m], far[eo : eo + s.m], resc[eo : eo + s.m],
                     lv[lo : lo + counts[i]], it[lo : lo + counts[i]], counts[i], seg, stats.refine_lanes,
                     windows=bool(stats.scan_windows))
        eo += s.m
        lo += counts[i]


@pytest.mark.parametrize("windows", [0, 1])
def test_brackets_match_oracle(capi, windows):
    """The bracket pass (epseon_get_brackets) against the oracle's
    bracket_level from the same scan arrays: ends, node counts, floor flag
    and band bit-exact; the ends straddle each found level's node count."""
    specs = batched(4096)[:40]
    _log(f"case: brackets of the first 40 batched potentials, windows {windows}")
    counts = [s.bound_levels() + 2 for s in specs]
    tables = [s.table() for s in specs]
    with capi.Context(0) as ctx:
        t = ctx.tuning()
        t.refine_mode = capi.REFINE_SEGMENTED
        t.windows = windows
        ctx.set_tuning(t)
        ctx.load([_pot(capi, s, tb, c) for s, tb, c in zip(specs, tables, counts)])
        ctx.solve(1e-10, 200)
        nodes, far, resc = ctx.scan_results()
        br = ctx.brackets()
        flags = [ctx.potential_info(i).flags for i in range(len(specs))]
    assert len(br) == sum(counts)
    eo = lo = 0
    n_found = 0
    for i, s in enumerate(specs):
        mine = br[lo : lo + counts[i]]
        if flags[i] != 0:
            assert np.all(mine["found"] == 0)
        else:
            prob = problem_from_spec(s)
            prob.v = tables[i]
            win = prob.windows(bool(windows))
            fe = oracle.band_far_exps(win)
            nd, fr, rs = nodes[eo : eo + s.m], far[eo : eo + s.m], resc[eo : eo + s.m]
            for v in range(counts[i]):
                ref = oracle.bracket_level(nd, fr, rs, prob.e_lo, prob.e_top, v, fe, win)
                got = mine[v]
                assert int(got["found"]) == ref.found, (i, v)
                if not ref.found:
                    continue
                n_found += 1
                assert np.float64(got["e_hi"]).view(np.int64) == np.float64(ref.e_hi).view(np.int64), (i, v)
                assert np.float64(got["e_lo"]).view(np.int64) == np.float64(ref.e_lo).view(np.int64), (i, v)
                assert (int(got["nodes_lo"]), int(got["nodes_hi"]), int(got["lo_valid"]), int(got["band"])) == (
                    ref.c_lo, ref.c_hi, ref.lo_valid, ref.band), (i, v)
                assert got["nodes_lo"] <= v < got["nodes_hi"] and got["e_lo"] < got["e_hi"]
        eo += s.m
        lo += counts[i]
    assert n_found > 1000


@pytest.mark.parametrize("mode", ["lane", "segmented"])
def test_every_node_stride_matches_oracle(capi, mode):
    """Two batch potentials of each node-sampling stride (1, 2, 4, 8, 16 and
    32 steps between sign samples; the counter packs the samples' sign bits
    and flushes them on a stride-dependent schedule) bit-exact against the
    oracle, whole-grid lane shots or windowed segmented shots."""
    specs = batched(4096)
    picked: dict[int, list[int]] = {}
    for i, s in enumerate(specs):
        prob = problem_from_spec(s)
        if prob.usable and len(picked.setdefault(prob.stride, [])) < 2:
            picked[prob.stride].append(i)
        if len(picked) == 6 and all(len(v) == 2 for v in picked.values()):
            break
    assert sorted(picked) == [1, 2, 4, 8, 16, 32]
    idx = [i for b in sorted(picked) for i in picked[b]]
    sub = [specs[i] for i in idx]
    _log(f"case: {mode}, potentials {idx} (two per stride)")
    counts = [s.bound_levels() + 2 for s in sub]
    tables = [s.table() for s in sub]
    seg = mode == "segmented"
    with capi.Context(0) as ctx:
        t = ctx.tuning()
        t.refine_mode = capi.REFINE_SEGMENTED if seg else capi.REFINE_LANE
        t.levels_per_lane = 1
        t.windows = 1 if seg else 0
        t.segmented_lanes = 64
        t.solve_split = 0.0
        ctx.set_tuning(t)
        ctx.load([_pot(capi, s, tb, c) for s, tb, c in zip(sub, tables, counts)])
        ctx.solve(1e-10, 200)
        nodes, far, resc = ctx.scan_results()
        lv, st, it = ctx.levels()
        lanes = ctx.stats().refine_lanes
    eo = lo = 0
    for i, s in enumerate(sub):
        prob = problem_from_spec(s)
        prob.v = tables[i]
        _compare(prob, nodes[eo : eo + s.m], far[eo : eo + s.m], resc[eo : eo + s.m],
                 lv[lo : lo + counts[i]], it[lo : lo + counts[i]], counts[i], seg, lanes)
        eo += s.m
        lo += counts[i]
