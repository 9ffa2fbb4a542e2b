"""GPU parity: the HIP path (through the C ABI) against the C oracle.

Node counts, far-end values, rescale counts and refined levels must be
bit-identical: both sides run the same FP64 operations in the same order
with explicit fma and no contraction.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from epseon_backend_amd.workloads import batched, double_wells
from tests.helpers import CFG_H2, morse_problem, problem_from_spec, random_morse_batch

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


def _to_capi(capi, prob: oracle.Problem, level_min: int = 0, level_count: int = 1):
    return capi.Potential(
        values=prob.v,
        r_min=prob.r_min,
        r_max=prob.r_max,
        reduced_mass=prob.mu,
        e_top=prob.e_top,
        n_energies=prob.m,
        level_min=level_min,
        level_count=level_count,
    )


def _set_mode(ctx, capi, seg: bool) -> None:
    """Lane mode over the whole grid, or segmented mode with band windows
    (forced on: these batches are too small for the automatic choice)."""
    t = ctx.tuning()
    t.refine_mode = capi.REFINE_SEGMENTED if seg else capi.REFINE_LANE
    t.windows = int(seg)
    ctx.set_tuning(t)


def test_device_is_gfx950(capi):
    info = capi.device_info(0)
    _log(f"device: {info.name.decode()} {info.arch.decode()} CUs={info.compute_units}")
    assert info.arch.decode().startswith("gfx950")


@pytest.mark.parametrize("windows", [True, False])
def test_scan_parity_h2(capi, windows):
    _log(f"case: scan parity, H2 Morse N=10000 M=2048, windows {windows}")
    prob = morse_problem(**CFG_H2)
    with capi.Context(0) as ctx:
        t = ctx.tuning()
        t.windows = int(windows)
        ctx.set_tuning(t)
        ctx.load([_to_capi(capi, prob)])
        info = ctx.potential_info(0)
        assert info.beta == prob.beta
        assert info.stride == prob.stride
        assert info.e_lo == prob.e_lo
        ctx.scan()
        nodes, far, resc = ctx.scan_results()
        lo, hi, fe = ctx.band_windows(0)
    win = prob.windows(windows)
    np.testing.assert_array_equal(lo, win[:, 0])
    np.testing.assert_array_equal(hi, win[:, 1])
    np.testing.assert_array_equal(fe, oracle.band_far_exps(win))
    on, of, orr = prob.scan(windows)
    np.testing.assert_array_equal(nodes, on)
    np.testing.assert_array_equal(resc, orr)
    np.testing.assert_array_equal(far.view(np.int64), of.view(np.int64))
    assert nodes.max() == 17


def test_band_windows_match_oracle(capi):
    """The device's band windows (fixed-point depth sums in wavefront prefix
    sums) equal the oracle's sequential ones, for a batch of uneven grids."""
    _log("case: band windows of 40 batched potentials and 2 double wells vs the oracle")
    specs = batched(4096)[:40] + double_wells(64, n_points=200000, n_energies=4096)[:2]
    probs = [problem_from_spec(s) for s in specs]
    with capi.Context(0) as ctx:
        t = ctx.tuning()
        t.windows = 1
        ctx.set_tuning(t)
        ctx.load([_to_capi(capi, p, 0, 4) for p in probs])
        got = [ctx.band_windows(i) for i in range(len(probs))]
        t.windows = 0
        ctx.set_tuning(t)
        whole = [ctx.band_windows(i) for i in range(len(probs))]
    for p, (lo, hi, fe) in zip(probs, whole):
        np.testing.assert_array_equal(lo, 0)
        np.testing.assert_array_equal(hi, p.trimmed[1] - p.trimmed[0])
    for p, (lo, hi, fe) in zip(probs, got):
        if not p.usable:
            continue
        win = p.windows()
        np.testing.assert_array_equal(lo, win[:, 0])
        np.testing.assert_array_equal(hi, win[:, 1])
        np.testing.assert_array_equal(fe, oracle.band_far_exps(win))


@pytest.mark.parametrize("mode", ["lane", "segmented"])
def test_solve_parity_h2(capi, mode):
    _log(f"case: solve parity ({mode} refine), H2 Morse, levels 0..19")
    prob = morse_problem(**CFG_H2)
    seg = mode == "segmented"
    with capi.Context(0) as ctx:
        _set_mode(ctx, capi, seg)
        ctx.load([_to_capi(capi, prob, 0, 20)])
        ctx.solve(1e-10, 200)
        lv, st, it = ctx.levels()
        stats = ctx.stats()
        assert stats.refine_mode == (capi.REFINE_SEGMENTED if seg else capi.REFINE_LANE)
        assert stats.scan_windows == int(seg)
    _log(f"  refine lanes per level: {stats.refine_lanes}")
    # Lane mode shoots the whole grid; segmented mode the band windows.
    _, _, _, olv, oit = oracle.solve(prob, 0, 20, 1e-10, 200, segmented=seg, lanes=stats.refine_lanes,
                                     windows=seg)
    np.testing.assert_array_equal(st[:17], 0)
    np.testing.assert_array_equal(st[17:], 1)
    np.testing.assert_array_equal(lv[:17].view(np.int64), olv[:17].view(np.int64))
    np.testing.assert_array_equal(it[:17], oit[:17])
    assert np.all(np.isnan(lv[17:]))
    # Physics sanity only (the oracle tests pin the analytic agreement): the
    # hard wall at r_min = 0.2 A lifts excited levels by ~1e-8 relative.
    ana = oracle.morse_levels(prob.mu, CFG_H2["de"], CFG_H2["a"], np.arange(10))
    np.testing.assert_allclose(lv[:10], ana, rtol=1e-7)


@pytest.mark.parametrize("mode", ["lane", "segmented"])
def test_ragged_batch_parity(capi, mode):
    _log(f"case: ragged batch of 12 random Morse potentials (seed 7), {mode} refine")
    probs = random_morse_batch(12, seed=7, n_range=(300, 6000), m_range=(40, 700))
    seg = mode == "segmented"
    with capi.Context(0) as ctx:
        _set_mode(ctx, capi, seg)
        ctx.load([_to_capi(capi, p, 0, 30) for p in probs])
        ctx.solve(1e-10, 200)
        nodes, far, resc = ctx.scan_results()
        lv, st, it = ctx.levels()
        strides = [ctx.potential_info(i).stride for i in range(len(probs))]
        lanes = ctx.stats().refine_lanes
        flags = [ctx.potential_info(i).flags for i in range(len(probs))]
    eo = lo = 0
    for i, p in enumerate(probs):
        assert strides[i] == p.stride
        assert flags[i] == (0 if p.usable else 1)
        if flags[i] != 0:
            assert np.all(st[lo : lo + 30] == 3)
            eo += p.m
            lo += 30
            continue
        on, of, orr, olv, oit = oracle.solve(p, 0, 30, 1e-10, 200, segmented=seg, lanes=lanes, windows=seg)
        np.testing.assert_array_equal(nodes[eo : eo + p.m], on)
        np.testing.assert_array_equal(resc[eo : eo + p.m], orr)
        np.testing.assert_array_equal(far[eo : eo + p.m].view(np.int64), of.view(np.int64))
        got = lv[lo : lo + 30]
        ok = ~np.isnan(olv)
        np.testing.assert_array_equal(np.isnan(got), ~ok)
        np.testing.assert_array_equal(got[ok].view(np.int64), olv[ok].view(np.int64))
        np.testing.assert_array_equal(it[lo : lo + 30][ok], oit[ok])
        eo += p.m
        lo += 30


def test_device_load_matches_host_load(capi):
    """epseon_load_potentials_device (V already in device memory, statistics
    and wall trimming on the GPU) gives bit-identical derived values and
    results to the host upload of the same tables, and reports a non-finite
    value with the host loader's message."""
    import torch

    from epseon_backend_amd.workloads import batched

    _log("case: 60 batched potentials + a wall-trimmed potential, host vs device upload")
    specs = batched(4096)[:60]
    pots = [capi.Potential(values=s.table(), r_min=s.r_min, r_max=s.r_max, reduced_mass=s.mu, e_top=s.de,
                           n_energies=s.m, level_min=0, level_count=s.bound_levels() + 2) for s in specs]
    # The reference test's potential (D_e 5500, a 10, [0, 10] A): V(0) ~ 9e8 cm^-1, trimmed.
    wall = morse_problem(mu=43.81, de=5500.0, a=10.0, re=0.6, r_min=0.0, r_max=10.0, n=16500, m=512)
    pots.append(capi.Potential(values=wall.v, r_min=0.0, r_max=10.0, reduced_mass=43.81, e_top=5500.0,
                               n_energies=512, level_min=0, level_count=4))
    offsets = np.cumsum([0] + [len(p.values) for p in pots])[:-1]
    flat = np.concatenate([p.values for p in pots])
    outs = []
    for device in (False, True):
        with capi.Context(0) as ctx:
            if device:
                v = torch.from_numpy(flat).cuda()
                ctx.load_device(pots, v.data_ptr(), offsets)
            else:
                ctx.load(pots)
            infos = [ctx.potential_info(i) for i in range(len(pots))]
            ctx.solve(1e-10, 200)
            nodes, far, resc = ctx.scan_results()
            lv, st, it = ctx.levels()
            outs.append(([(i.first_point, i.point_count, i.flags, i.stride, np.float64(i.e_lo).view(np.int64))
                          for i in infos], nodes, far.view(np.int64), resc, lv.view(np.int64), st, it))
    assert outs[0][0] == outs[1][0]
    assert outs[0][0][-1][0] > 0  # the wall potential was trimmed
    for a, b in zip(outs[0][1:], outs[1][1:]):
        np.testing.assert_array_equal(a, b)
    bad = flat.copy()
    bad[offsets[3] + 10] = np.nan
    with capi.Context(0) as ctx:
        v = torch.from_numpy(bad).cuda()
        with pytest.raises(capi.EpseonError, match=r"potential 3: V\[10\] is not finite"):
            ctx.load_device(pots, v.data_ptr(), offsets)
