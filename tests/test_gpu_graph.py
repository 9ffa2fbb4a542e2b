"""Graph replay of the solve (tuning `graph`): when a solve repeats the
previous one's batch, tuning and tolerance, the scan, bracket and refinement
launches are recorded as a HIP graph and replayed from then on.

* every replay equals the eager solve bit for bit, with eager work (a plain
  scan and refinement at another tolerance) and host synchronisation between
  the replays;
* a new tolerance, a new tuning or a new batch launches directly once, then
  records a new graph, whose replays again equal the eager solves;
* the first replays are checked against the oracle on a few potentials.
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


def _pots(capi, specs, counts):
    return [capi.Potential(values=s.table(), r_min=s.r_min, r_max=s.r_max, reduced_mass=s.mu, e_top=s.de,
                           n_energies=s.m, level_min=0, level_count=c) for s, c in zip(specs, counts)]


def _results(ctx):
    nodes, far, resc = ctx.scan_results()
    lv, st, it = ctx.levels()
    return nodes, far.view(np.int64), resc, lv.view(np.int64), st, it


def _equal(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def _tune(ctx, capi, seg: bool, windows: int, graph: int) -> None:
    t = ctx.tuning()
    t.refine_mode = capi.REFINE_SEGMENTED if seg else capi.REFINE_LANE
    t.levels_per_lane = 1
    t.windows = windows
    t.solve_split = 0.0
    t.graph = graph
    ctx.set_tuning(t)


@pytest.mark.parametrize("mode,windows", [("segmented", 1), ("lane", 0)])
def test_graph_replays_match_eager(capi, mode, windows):
    seg = mode == "segmented"
    specs = batched(4096)[:200]
    counts = [s.bound_levels() + 2 for s in specs]
    _log(f"case: first 200 batched potentials, {mode} refine, windows {windows}: graph replays vs eager")
    with capi.Context(0) as ctx:
        _tune(ctx, capi, seg, windows, graph=0)
        ctx.load(_pots(capi, specs, counts))
        ctx.solve(1e-10, 200)
        assert ctx.stats().solve_graph == 0
        eager = _results(ctx)
        ctx.solve(1e-12, 200)
        eager_tight = _results(ctx)
        _tune(ctx, capi, seg, windows, graph=1)
        for rep in range(5):
            ctx.solve(1e-10, 200)
            stats = ctx.stats()
            # The previous solve (1e-12) had another key: the first launches
            # directly, the next ones record and replay.
            assert stats.solve_graph == (1 if rep > 0 else 0) and stats.scan_windows == windows
            _equal(_results(ctx), eager)
            _log(f"  solve {rep} (graph {stats.solve_graph}): {stats.solve_ms:.3f} ms, identical")
            # Eager work on the same buffers between the replays.
            ctx.scan()
            ctx.refine(1e-12, 200)
            _equal(_results(ctx), eager_tight)
        # Another tolerance: direct once, then a new graph; going back records again.
        for tol, want in ((1e-12, eager_tight), (1e-10, eager), (1e-12, eager_tight)):
            for rep in range(3):
                ctx.solve(tol, 200)
                assert ctx.stats().solve_graph == (1 if rep > 0 else 0)
                _equal(_results(ctx), want)
        # Another scan workgroup size: a new graph, the same results.
        t = ctx.tuning()
        t.scan_waves_per_block = 2
        ctx.set_tuning(t)
        for rep in range(2):
            ctx.solve(1e-10, 200)
            assert ctx.stats().solve_graph == rep
            _equal(_results(ctx), eager)
        lanes = ctx.stats().refine_lanes
    # The oracle agrees on the first potentials.
    nodes, far, resc, lvi, st, it = eager
    lv = lvi.view(np.float64)
    eo = lo = 0
    for s, c in zip(specs[:12], counts[:12]):
        prob = problem_from_spec(s)
        if prob.usable:
            on, of, orr, olv, oit = oracle.solve(prob, 0, c, 1e-10, 200, segmented=seg, lanes=lanes,
                                                 windows=bool(windows))
            np.testing.assert_array_equal(nodes[eo : eo + s.m], on)
            np.testing.assert_array_equal(far[eo : eo + s.m], of.view(np.int64))
            ok = ~np.isnan(olv)
            np.testing.assert_array_equal(lv[lo : lo + c][ok].view(np.int64), olv[ok].view(np.int64))
            np.testing.assert_array_equal(it[lo : lo + c][ok], oit[ok])
        eo += s.m
        lo += c


def test_graph_follows_reloads(capi):
    """A load drops the recorded graph (its buffers are reused): replays after
    each reload equal an eager solve of the same batch on another context."""
    specs = batched(4096)
    batches = [specs[200:260], specs[1000:1100], specs[200:260]]
    _log("case: three batches loaded in turn on one context with graph replay")
    with capi.Context(0) as ctx, capi.Context(0) as ref:
        _tune(ctx, capi, True, 1, graph=1)
        _tune(ref, capi, True, 1, graph=0)
        for b in batches:
            counts = [s.bound_levels() + 2 for s in b]
            ctx.load(_pots(capi, b, counts))
            ref.load(_pots(capi, b, counts))
            ref.solve(1e-10, 200)
            want = _results(ref)
            for rep in range(3):
                ctx.solve(1e-10, 200)
                assert ctx.stats().solve_graph == (1 if rep > 0 else 0)
                _equal(_results(ctx), want)


def test_levels_into_reused_arrays_and_summary(capi):
    """levels(out=...) fills caller arrays (larger ones keep their tail) with
    the same values as fresh ones; the statistics' device summary counts the
    converged levels and matches between graph replays and eager solves."""
    specs = batched(4096)[300:340]
    counts = [s.bound_levels() + 2 for s in specs]
    n = sum(counts)
    _log(f"case: {len(specs)} potentials, {n} levels, reused result arrays")
    with capi.Context(0) as ctx, capi.Context(0) as ref:
        _tune(ctx, capi, True, 1, graph=1)
        _tune(ref, capi, True, 1, graph=0)
        ctx.load(_pots(capi, specs, counts))
        ref.load(_pots(capi, specs, counts))
        ref.solve(1e-10, 200)
        want = ref.levels()
        want_stats = ref.stats()
        out = (np.full(n + 5, -1.0), np.full(n + 5, -7, np.int32), np.full(n + 5, -9, np.int32))
        for rep in range(3):
            ctx.solve(1e-10, 200)
            got = ctx.levels(out)
            for g, w in zip(got, want):
                np.testing.assert_array_equal(np.asarray(g)[:n].view(np.int64 if g.dtype == np.float64 else np.int32),
                                              np.asarray(w).view(np.int64 if w.dtype == np.float64 else np.int32))
            assert np.all(out[0][n:] == -1.0) and np.all(out[1][n:] == -7) and np.all(out[2][n:] == -9)
            s = ctx.stats()
            assert s.levels_converged == int(np.sum(want[1] == capi.LEVEL_OK))
            assert s.refine_point_updates == want_stats.refine_point_updates > 0
            assert s.levels_converged == want_stats.levels_converged
        # The summary restated on the host: iterations x window steps of every
        # level that shot (segmented shots walk their band's window), and the
        # trimmed grid's steps of every converged level.
        br = ctx.brackets()
        _, st, it = want
        pu = 0
        shot = 0
        o = 0
        for p, c in enumerate(counts):
            lo, hi, _ = ctx.band_windows(p)
            info = ctx.potential_info(p)
            for l in range(c):
                if it[o + l] > 0:
                    pu += int(it[o + l]) * int(hi[br["band"][o + l]] - lo[br["band"][o + l]] - 1)
                if st[o + l] == capi.LEVEL_OK:
                    shot += int(info.point_count) - 2
            o += c
        assert want_stats.refine_point_updates == pu
        ctx.normalize()
        assert ctx.stats().normalize_point_updates == shot
        with pytest.raises(ValueError):
            ctx.levels((np.empty(n), np.empty(n, np.int64), np.empty(n, np.int32)))
        with pytest.raises(ValueError):
            ctx.levels((np.empty(n - 1), np.empty(n, np.int32), np.empty(n, np.int32)))
