"""Pin the C oracle (oracle/numerov_oracle.c) before trusting it.

The reference ships no numerics for this path and no golden vectors, so each
oracle function is checked against something the mathematics fixes:

* constant potential: the discrete solution is y_n = sin(n theta)/sin(theta)
  with 2 cos(theta) = (12 - 10 f)/f, and the Dirichlet levels of n intervals
  are E_k = (12 K'/h^2)(1 - cos(k pi/n))/(5 + cos(k pi/n)) -- trajectory, node
  counts and refined levels;
* a dense symmetric eigensolve of the same discrete problem,
  (diag(V) - K' B^-1 A) y = E y with A = tridiag(1,-2,1)/h^2 and
  B = tridiag(1,10,1)/12 -- levels and Sturm counts;
* 50-digit mpmath trajectories of the textbook (division) form of Numerov's
  method, committed in tests/golden/numerov_trajectories.json;
* analytic Morse levels (tests/golden/morse_levels.json) and the h^4
  convergence of the scheme (Richardson h vs h/2).

A dropped term, a wrong coefficient, a sign or an index slip in the
recurrence fails the closed forms or the eigensolve.
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import mpmath as mp
import numpy as np
import pytest
import scipy.linalg

import oracle
from tests.helpers import CFG_H2, CFG_I2, morse_problem, random_morse_batch

GOLDEN = Path(__file__).resolve().parent / "golden"
K = oracle.K_CM


def box_problem(n: int, mu: float, length: float, e_top: float, m: int) -> oracle.Problem:
    return oracle.Problem(v=np.zeros(n), r_min=0.0, r_max=length, mu=mu, e_top=e_top, m=m)


def box_levels(n_points: int, h: float, mu: float, count: int) -> np.ndarray:
    n = n_points - 1  # intervals
    kp = K / mu
    k = np.arange(1, count + 1)
    x = k * np.pi / n
    one_minus_c = 2.0 * np.sin(0.5 * x) ** 2  # 1 - cos x without cancellation
    return (12.0 * kp / h**2) * one_minus_c / (6.0 - one_minus_c)


# ------------------------------------------------------------ closed forms


def test_constant_potential_trajectory():
    n, mu, length = 200, 1.0, 1.0
    prob = box_problem(n, mu, length, e_top=1.0, m=1)
    b = prob.beta
    for energy in (50.0, 700.0, 3000.0):
        e = b * energy
        w = oracle.trajectory(prob.ut, e)
        mp.mp.dps = 40
        f = 1 + mp.mpf(b) * mp.mpf(energy)
        theta = mp.acos((12 / f - 10) / 2)
        ref = np.array([float(f ** (i - 1) * mp.sin(i * theta) / mp.sin(theta)) for i in range(n)])
        np.testing.assert_allclose(w[1:], ref[1:], rtol=2e-11, atol=1e-12 * np.abs(ref).max())
        assert w[0] == 0.0


def test_box_levels_and_counts():
    n, mu, length = 4001, 2.0, 3.0
    prob = box_problem(n, mu, length, e_top=400.0, m=700)
    ref = box_levels(n, prob.h, mu, 60)
    nbound = int(np.sum(ref < prob.e_top))
    for seg in (False, True):
        nodes, _, _, lv, _ = oracle.solve(prob, 0, nbound, 1e-12, 300, segmented=seg, lanes=128)
        np.testing.assert_allclose(lv, ref[:nbound], rtol=5e-12)
    # Sturm: the node count at E is the number of levels below E.
    energies = oracle.trial_energies(prob.e_lo, prob.e_top, prob.m)
    expect = np.searchsorted(ref, energies)
    np.testing.assert_array_equal(nodes, expect)


def test_box_diff_form_far_end_sign():
    """The far-end value changes sign exactly at the closed-form levels."""
    n, mu, length = 1201, 0.7, 2.0
    prob = box_problem(n, mu, length, e_top=2000.0, m=1)
    ref = box_levels(n, prob.h, mu, 10)
    for k, ek in enumerate(ref):
        below = oracle.shoot_diff(prob.uv, prob.beta * ek * (1 - 1e-9), prob.stride)
        above = oracle.shoot_diff(prob.uv, prob.beta * ek * (1 + 1e-9), prob.stride)
        assert below[0] == k and above[0] == k + 1
        assert math.copysign(1, below[1]) != math.copysign(1, above[1])


# -------------------------------------------------------- dense eigensolve


def discrete_hamiltonian(v: np.ndarray, h: float, mu: float) -> np.ndarray:
    m = len(v) - 2
    a = (np.diag(np.full(m, -2.0)) + np.diag(np.ones(m - 1), 1) + np.diag(np.ones(m - 1), -1)) / h**2
    bm = (np.diag(np.full(m, 10.0)) + np.diag(np.ones(m - 1), 1) + np.diag(np.ones(m - 1), -1)) / 12.0
    x = scipy.linalg.solve(bm, a, assume_a="sym")
    ham = np.diag(v[1:-1]) - (K / mu) * x
    return 0.5 * (ham + ham.T)


@pytest.fixture(scope="module")
def small_morse():
    prob = morse_problem(mu=0.50391, de=38297.0, a=1.9426, re=0.7414, r_min=0.35, r_max=2.5, n=700, m=300)
    first, last, ok = prob.trimmed
    assert (first, last, ok) == (0, 699, True)
    eig = scipy.linalg.eigvalsh(discrete_hamiltonian(prob.v, prob.h, prob.mu))
    return prob, eig


def test_levels_match_dense_eigensolve(small_morse):
    prob, eig = small_morse
    nb = int(np.sum(eig < prob.e_top))
    for seg in (False, True):
        _, _, _, lv, it = oracle.solve(prob, 0, nb + 2, 1e-12, 300, segmented=seg, lanes=64)
        np.testing.assert_allclose(lv[:nb], eig[:nb], rtol=3e-11)
        assert np.all(np.isnan(lv[nb:]))


def test_sturm_counts_match_dense_eigensolve(small_morse):
    prob, eig = small_morse
    rng = np.random.default_rng(5)
    for energy in rng.uniform(prob.e_lo, prob.e_top, 200):
        expect = int(np.sum(eig < energy))
        e = prob.beta * energy
        assert oracle.shoot(prob.pt, e, prob.stride)[0] == expect
        assert oracle.shoot_diff(prob.uv, e, prob.stride)[0] == expect
        assert oracle.shoot_segmented(prob.uv, e, prob.stride, lanes=16)[2] == expect


# ------------------------------------------------------------- mpmath golden


@pytest.fixture(scope="module")
def golden_trajectories():
    return json.loads((GOLDEN / "numerov_trajectories.json").read_text())


# Golden energies placed on a level of their short grid: there the tail is a
# cancellation and input rounding is amplified by ~1e5-1e6.
NEAR_LEVEL = {"2167.25", "107.135"}


def test_trajectories_match_mpmath(golden_trajectories):
    for case in golden_trajectories["cases"]:
        v = np.array([float(x) for x in case["v"]])
        y_ref = np.array([float(x) for x in case["y"]])
        prob = oracle.Problem(v=v, r_min=float(case["r_min"]), r_max=float(case["r_max"]),
                              mu=float(case["mu"]), e_top=float(case["de"]), m=1)
        e = prob.beta * float(case["energy"])
        ut = oracle.shifted_potential(v, prob.beta)
        w = oracle.trajectory(ut, e)
        f = e - ut
        # y_i = w_i / (f_i prod_{j<i} f_j); the oracle starts from w_1 = 1,
        # the golden from y_1 = 1.
        p = np.cumprod(f)
        y = w / p
        y1 = y[1]
        y = y / y1
        rel5 = np.max(np.abs(y[1:] - y_ref[1:]) / np.abs(y_ref[1:]))
        # difference form: same recurrence, far-end value
        c, wd, k = oracle.shoot_diff(oracle.scaled_potential(v, prob.beta), e, 1)
        assert k == 0
        yd = wd / p[-1] / y1
        reld = abs(yd - y_ref[-1]) / abs(y_ref[-1])
        if case["energy"] in NEAR_LEVEL:
            assert rel5 < 1e-6 and reld < 1e-8, (case["energy"], rel5, reld)
            assert reld < rel5 / 10, (rel5, reld)
        else:
            assert rel5 < 1e-11, (case["energy"], rel5)
            assert reld < 1e-13, (case["energy"], reld)


# ------------------------------------------------------------ physics checks


def test_analytic_morse_i2():
    golden = [float(x) for x in json.loads((GOLDEN / "morse_levels.json").read_text())["i2"]]
    prob = morse_problem(**dict(CFG_I2, m=2048))
    _, _, _, lv, _ = oracle.solve(prob, 0, 100, 1e-12, 300)
    # Numerov error ~ (k h)^4/240 < 1e-11 here; the top ~15 levels feel the box.
    np.testing.assert_allclose(lv, golden[:100], rtol=2e-11)


def test_analytic_morse_h2_low_levels():
    golden = [float(x) for x in json.loads((GOLDEN / "morse_levels.json").read_text())["h2"]]
    prob = morse_problem(**CFG_H2)
    _, _, _, lv, _ = oracle.solve(prob, 0, 17, 1e-12, 300)
    assert np.sum(~np.isnan(lv)) == 17
    # Low levels: discretisation error only; higher ones feel the wall at 0.2 A.
    np.testing.assert_allclose(lv[:3], golden[:3], rtol=1e-10)
    np.testing.assert_allclose(lv[:12], golden[:12], rtol=1e-7)


def test_fourth_order_convergence():
    """Halving h cuts the level error by ~16 (Numerov is O(h^4))."""
    de, re, a, mu = 38297.0, 0.7414, 1.9426, 0.50391
    errs = []
    golden = [float(x) for x in json.loads((GOLDEN / "morse_levels.json").read_text())["h2"]]
    for n in (501, 1001, 2001):
        prob = morse_problem(mu=mu, de=de, a=a, re=re, r_min=0.1, r_max=3.0, n=n, m=200)
        _, _, _, lv, _ = oracle.solve(prob, 0, 4, 1e-13, 300)
        errs.append(np.abs(lv - golden[:4]))
    ratio1 = errs[0] / errs[1]
    ratio2 = errs[1] / errs[2]
    assert np.all((ratio1 > 13) & (ratio1 < 19)), ratio1
    assert np.all((ratio2 > 13) & (ratio2 < 19)), ratio2


# --------------------------------------------------------- internal rules


@pytest.mark.parametrize("seed", [3, 11])
def test_sampled_count_equals_full_count(seed):
    probs = random_morse_batch(6, seed=seed, n_range=(400, 5000), m_range=(10, 20))
    rng = np.random.default_rng(seed)
    strides = set()
    for p in probs:
        if not p.usable:
            continue
        strides.add(p.stride)
        for energy in rng.uniform(p.e_lo, p.e_top, 60):
            e = p.beta * energy
            assert oracle.shoot(p.pt, e, p.stride)[0] == oracle.shoot_full_count(p.ut, e)
    assert len(strides) >= 2


def test_node_stride_rule():
    for mu, h in ((0.5, 1e-3), (63.0, 3e-5), (20.0, 5e-3), (100.0, 2e-2)):
        b = h * h * mu / (12 * K)
        for e_top in (1e3, 1e4, 5e4):
            s = oracle.node_stride(b, e_top, 0.0)
            f = 1 + b * e_top
            u = 12 / f - 10
            theta = math.acos(u / 2) if u > -2 else math.pi
            assert s in (1, 2, 4, 8, 16, 32)
            assert s == 1 or s * theta <= math.pi / 2
            assert s == 32 or 2 * s * theta > math.pi / 2


def test_wall_trimming_of_reference_test_potential():
    """The reference's own test potential (D_e 5500, r_e 0.6, a 10 on [0, 10],
    16500 points, test_libepseon_gpu.py:185-196) reaches 9e8 cm^-1 at r = 0;
    the Dirichlet point moves to the first point below V_min + 1/(2 beta)."""
    prob = morse_problem(mu=87.62 / 2, de=5500.0, a=10.0, re=0.6, r_min=0.0, r_max=10.0, n=16500, m=256)
    first, last, ok = prob.trimmed
    assert ok and last == 16499
    wall = prob.e_lo + 0.5 / prob.beta
    assert prob.v[first] >= wall and prob.v[first + 1] < wall
    f = 1 + prob.beta * (prob.e_lo - prob.v[first + 1 : last])
    assert f.min() > 0.5


def test_bracket_invariants():
    prob = morse_problem(**CFG_H2)
    nodes, far, resc = oracle.scan(prob.pt, prob.beta, prob.e_lo, prob.e_top, prob.m, prob.stride)
    assert np.all(np.diff(nodes) >= 0)
    for v in range(17):
        br = oracle.bracket_level(nodes, far, resc, prob.e_lo, prob.e_top, v, prob.far_exp)
        assert br.found
        assert br.c_lo <= v < br.c_hi
        assert br.e_lo < br.e_hi
        lv, it = oracle.refine_level(prob.uv, prob.beta, prob.stride, v, br, 1e-10, prob.abs_tol)
        assert br.e_lo < lv < br.e_hi and 0 < it < 40
    br = oracle.bracket_level(nodes, far, resc, prob.e_lo, prob.e_top, 17, prob.far_exp)
    assert not br.found


def test_segmented_and_sequential_refine_agree():
    prob = morse_problem(**dict(CFG_I2, m=1024))
    a = oracle.solve(prob, 0, 110, 1e-11, 300, segmented=False)[3]
    for lanes in (64, 256):
        b = oracle.solve(prob, 0, 110, 1e-11, 300, segmented=True, lanes=lanes)[3]
        ok = ~np.isnan(a)
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
        np.testing.assert_allclose(b[ok], a[ok], rtol=1e-11)


def test_scan_form_far_end_matches_w_form():
    """The 12-scaled scan shot (v_i = w_i / 12^(i-1), 4 operations per step)
    reports w_{N-1} = mantissa 2^(I - 480 r): against the unscaled
    division-free trajectory on a short grid (no rescales in either) and
    against the difference-form shot (which rescales the other way) on a
    long grid where the scan form scales up many times."""
    prob = morse_problem(mu=7.0, de=20000.0, a=2.0, re=1.2, r_min=0.7, r_max=3.5, n=420, m=64)
    assert prob.usable
    rng = np.random.default_rng(4)
    for energy in rng.uniform(prob.e_lo + 50, prob.e_top, 20):
        e = prob.beta * energy
        ref = oracle.trajectory(prob.ut, e)[-1]
        np.testing.assert_allclose(oracle.scan_far_end(prob.pt, e, prob.stride), ref, rtol=1e-11)
    big = morse_problem(mu=63.452, de=12440.0, a=1.8676, re=2.666, r_min=2.0, r_max=8.0, n=200_000, m=64)
    c, i = oracle.far_scale(len(big.pt))
    assert 1.0 <= c < 2.0 and i > 480
    for energy in (500.0, 3000.0, 7777.0, 12000.0):
        e = big.beta * energy
        cs, ws, rs = oracle.shoot(big.pt, e, big.stride)
        cd, wd, kd = oracle.shoot_diff(big.uv, e, big.stride)
        assert cs == cd and rs > 0
        # log2 |F| and sign agree (F = w 2^k in both forms).
        assert np.sign(ws) == np.sign(wd)
        ls = np.log2(abs(ws)) + i - 480 * rs
        ld = np.log2(abs(wd)) + 480 * kd
        assert abs(ls - ld) < 1e-6, (energy, ls, ld)


def test_refinement_model_elementary_functions():
    """The deterministic exp / log / sqrt that steer the refinement probes
    (written with basic IEEE operations so the kernels match bit for bit) are
    accurate to a few ulps over the ranges they see."""
    mexp, mlog, msqrt = oracle.model_functions()
    rng = np.random.default_rng(9)
    for y in np.concatenate([rng.uniform(-700, 700, 300), rng.uniform(-1, 1, 300), [0.0]]):
        assert abs(mexp(y) / math.exp(y) - 1) < 4e-15, y
    for x in np.concatenate([np.exp(rng.uniform(-600, 600, 300)), rng.uniform(0.5, 2, 300)]):
        assert abs(mlog(x) - math.log(x)) < 2e-15 * max(1.0, abs(math.log(x))), x
    for d in np.concatenate([rng.uniform(0, 4, 300), np.exp(rng.uniform(-200, 200, 300))]):
        assert (abs(msqrt(d) / math.sqrt(d) - 1) < 4.5e-16) if d > 0 else (msqrt(d) == 0.0), d
