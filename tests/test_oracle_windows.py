"""Shot windows of the oracle (numerov_oracle.c, windows section).

A shot stops where it has gone at least 15 e-folds into the classically
forbidden region past its outermost turning points (on both sides), and the
scan and the refinement use per-energy-band windows.  Pinned here:

- the growth bound g(x) <= acosh(1 + x/2) <= ln(lambda) the window depth is
  summed from (closed forms, dense x);
- the window rule itself (allowed region inside, depth reached exactly at the
  ends) and the nesting of the band windows;
- that a Dirichlet point at the window ends leaves the discrete eigenvalues
  unchanged: dense eigensolves of the windowed and the whole-grid problems;
- that windowed solves reproduce whole-grid solves: node counts bit-identical,
  levels within the refinement tolerance, on a batched subset, double wells
  and levels whose bracket floor comes from another band's window.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg

import oracle
from epseon_backend_amd.workloads import batched, double_wells
from tests.helpers import morse_problem, problem_from_spec

K = oracle.K_CM
DEPTH = 15.0


# floor(65536 acosh(1 + y/2)/sqrt(y)) at y = 2^k, k = -5..3.
RATIO = np.array([65450, 65366, 65199, 64871, 64242, 63073, 61029, 57761, 53116], dtype=np.uint64)


def growth_fixed(x: np.ndarray) -> np.ndarray:
    """The oracle's growth bound g(x) in 2^-40 units (numerov_oracle.c,
    eo_growth_term), restated with numpy integers."""
    x = np.asarray(x, dtype=np.float64)
    m, k = np.frexp(x)
    k = k.astype(np.int64)
    t = np.floor(np.ldexp(m, 16)).astype(np.uint64)
    big = k > 3
    k = np.where(big, 3, k)
    t = np.where(big, np.uint64(0xFFFF), t)
    e = k + 64
    h = e >> 1
    n = t << ((e & 1) + 32).astype(np.uint64)
    r = np.floor(np.sqrt(n.astype(np.float64))).astype(np.uint64)
    for _ in range(2):
        r = np.where(r * r > n, r - np.uint64(1), r)
        r = np.where((r + np.uint64(1)) * (r + np.uint64(1)) <= n, r + np.uint64(1), r)
    up = np.clip(h - 16, 0, 63).astype(np.uint64)
    down = np.clip(16 - h, 0, 63).astype(np.uint64)
    sq = np.where(h >= 16, r << up, r >> down)
    f = RATIO[np.clip(k, -5, 3) + 5]
    g = (sq * f) >> np.uint64(16)
    return np.where((x > 0) & (k >= -64), g, 0).astype(np.int64)


def growth_bound(x: np.ndarray) -> np.ndarray:
    return np.ldexp(growth_fixed(x).astype(np.float64), -40)


def test_growth_ratio_table():
    """The table holds floor(65536 acosh(1 + y/2)/sqrt(y)) at y = 2^k."""
    import mpmath

    mpmath.mp.dps = 40
    for i, k in enumerate(range(-5, 4)):
        y = mpmath.mpf(2) ** k
        assert int(RATIO[i]) == int(mpmath.floor(65536 * mpmath.acosh(1 + y / 2) / mpmath.sqrt(y)))


def test_growth_bound_is_below_the_recurrence_growth():
    x = np.concatenate([np.geomspace(1e-30, 6.0, 400001, endpoint=False), 2.0 ** np.arange(-90, 3)])
    x = x[x < 6.0]
    g = growth_bound(x)
    a = 2.0 * np.arcsinh(0.5 * np.sqrt(x))  # acosh(1 + x/2), accurate for small x
    assert np.all(g <= a)
    # Within 9 % of it wherever the 2^-40 resolution allows (so the windows
    # stay tight); one ulp of x below each binade edge is the worst case.
    fine = x >= 2.0**-40
    assert np.all(g[fine] > 0.91 * a[fine])
    edges = np.nextafter(2.0 ** np.arange(-40, 3), 0.0)
    edges = edges[edges < 6.0]
    assert np.all(growth_bound(edges) > 0.91 * 2.0 * np.arcsinh(0.5 * np.sqrt(edges)))
    # Constant-f Numerov recurrence f y+ = (12 - 10 f) y - f y-, f = 1 - x/12:
    # its roots satisfy lambda + 1/lambda = 12/f - 10 = 2 + x/f >= 2 + x, so
    # ln(lambda) = acosh(1 + x/(2f)) >= acosh(1 + x/2).
    f = 1.0 - x / 12.0
    assert np.all(f > 0.5)
    assert np.all(2.0 * np.arcsinh(0.5 * np.sqrt(x / f)) >= a)


def _check_window(uv: np.ndarray, e: float) -> tuple[int, int]:
    lo, hi = oracle.window(uv, e)
    allowed = np.nonzero(uv <= e)[0]
    assert lo <= allowed[0] and allowed[-1] <= hi
    x = 12.0 * (uv - e)
    depth = int(DEPTH) << 40
    if hi < len(uv) - 1:
        right = growth_fixed(x[allowed[-1] + 1 : hi + 1])
        assert right.sum() >= depth > right[:-1].sum()
    if lo > 0:
        left = growth_fixed(x[lo : allowed[0]])
        assert left.sum() >= depth > left[1:].sum()
    return lo, hi


def test_window_rule_and_band_nesting():
    specs = batched(4096)
    for s in (specs[3], specs[1536], specs[2816]):
        prob = problem_from_spec(s)
        uv = prob.uv
        for frac in (0.01, 0.2, 0.5, 0.8, 0.97):
            _check_window(uv, prob.beta * (prob.e_lo + frac * (prob.e_top - prob.e_lo)))
        win = prob.windows()
        assert np.all(np.diff(win[:, 0]) <= 0) and np.all(np.diff(win[:, 1]) >= 0)
        for b, (lo, hi) in enumerate(win):
            assert (lo, hi) == _check_window(uv, prob.beta * oracle.band_top(b, prob.e_lo, prob.e_top))
        full = prob.windows(False)
        assert np.all(full == [0, len(uv) - 1])
    # Double well: both wells (and the barrier between them) stay inside.
    dw = problem_from_spec(double_wells(64, n_points=20000, n_energies=2048)[0])
    for frac in (0.05, 0.3, 0.6):
        _check_window(dw.uv, dw.beta * (dw.e_lo + frac * (dw.e_top - dw.e_lo)))


def test_group_band_rule():
    e_lo, e_hi, m = 10.0, 5000.0, 300
    de = oracle.trial_spacing(e_lo, e_hi, m)
    tops = [oracle.band_top(b, e_lo, e_hi) for b in range(oracle.BANDS)]
    for j in range(m):
        jt = min(m - 1, (j // 64) * 64 + 63)
        et = oracle.trial_energies(e_lo, e_hi, m)[jt]
        expect = next((b for b in range(oracle.BANDS - 1) if tops[b] >= et), oracle.BANDS - 1)
        assert oracle.group_band(j, m, e_lo, e_hi) == expect
    assert de > 0


def _dense_levels(v: np.ndarray, h: float, mu: float) -> np.ndarray:
    m = len(v) - 2
    a = (np.diag(np.full(m, -2.0)) + np.diag(np.ones(m - 1), 1) + np.diag(np.ones(m - 1), -1)) / h**2
    bm = (np.diag(np.full(m, 10.0)) + np.diag(np.ones(m - 1), 1) + np.diag(np.ones(m - 1), -1)) / 12.0
    x = scipy.linalg.solve(bm, a, assume_a="sym")
    ham = np.diag(v[1:-1]) - (K / mu) * x
    return scipy.linalg.eigvalsh(0.5 * (ham + ham.T))


def test_window_leaves_dense_eigenvalues_unchanged():
    """Heavy Morse on a coarse grid: every band's window cuts deep into both
    walls, yet the levels below the band top agree with the whole-grid
    problem's to rounding (the truncation effect is ~e^-30 of a spacing)."""
    prob = morse_problem(mu=63.452, de=12440.0, a=1.8676, re=2.666, r_min=2.2, r_max=5.0, n=1400, m=512)
    first, last, ok = prob.trimmed
    assert ok and first == 0
    full = _dense_levels(prob.v, prob.h, prob.mu)
    win = prob.windows()
    assert win[0][1] - win[0][0] + 1 < prob.n // 2
    # Dense-solver rounding: a few ulp of the operator norm (~ the largest
    # eigenvalue).  The check has teeth: windows cut where the bound's sum
    # reaches 4 instead of 15 move these levels by ~2e-2 cm^-1 (8: ~6e-6, 10: ~1e-7;
    # e^-2 per unit of depth).
    atol = 64 * np.finfo(float).eps * np.abs(full).max()
    assert atol < 1e-8
    for b, (lo, hi) in enumerate(win):
        top = oracle.band_top(b, prob.e_lo, prob.e_top)
        part = _dense_levels(prob.v[lo : hi + 1], prob.h, prob.mu)
        k = int(np.sum(full < top))
        assert int(np.sum(part < top)) == k
        np.testing.assert_allclose(part[:k], full[:k], rtol=0, atol=atol)


def _compare_solves(prob, levels: int, segmented: bool, lanes: int = 64):
    a = oracle.solve(prob, 0, levels, 1e-10, 200, segmented=segmented, lanes=lanes, windows=True)
    b = oracle.solve(prob, 0, levels, 1e-10, 200, segmented=segmented, lanes=lanes, windows=False)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(np.isnan(a[3]), np.isnan(b[3]))
    ok = ~np.isnan(b[3])
    np.testing.assert_allclose(a[3][ok], b[3][ok], rtol=1e-10, atol=0)
    return a, b


@pytest.mark.parametrize("segmented", [False, True])
def test_windowed_solve_matches_whole_grid(segmented):
    specs = batched(4096)
    for i in (0, 5, 14, 47, 256, 1536):
        prob = problem_from_spec(specs[i])
        if prob.usable:
            _compare_solves(prob, specs[i].bound_levels() + 2, segmented)


def test_windowed_double_wells_match_whole_grid():
    for spec in double_wells(64, n_points=40000, n_energies=4096)[:3]:
        prob = problem_from_spec(spec)
        _compare_solves(prob, spec.bound_levels() + 2, True)


def test_foreign_floor_is_shot_again():
    """Levels whose bracket floor is a scan point of another band's window
    shoot that floor again in their own window first (one extra
    iteration), then converge to the whole-grid level."""
    specs = batched(4096)
    seen = 0
    for i, v in ((3, 56), (5, 13), (14, 55), (47, 7)):
        prob = problem_from_spec(specs[i])
        nodes, far, resc = prob.scan()
        br = oracle.bracket_level(nodes, far, resc, prob.e_lo, prob.e_top, v, prob.far_exps(), prob.windows())
        assert br.found and br.lo_foreign and br.rate == 0.0
        assert oracle.group_band(0, prob.m, prob.e_lo, prob.e_top) <= br.band
        lv, it = prob.refine(nodes, far, resc, v)
        whole = prob.refine(*prob.scan(False), v, on=False)
        assert it >= 2
        np.testing.assert_allclose(lv, whole[0], rtol=1e-10)
        seen += 1
    assert seen == 4
