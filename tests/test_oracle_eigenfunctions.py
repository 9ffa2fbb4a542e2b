"""Pin the oracle's eigenfunctions and adjoint (oracle/numerov_oracle.c,
eo_psi_norm / eo_psi / eo_level_vjp) before trusting them.

* constant potential: the discrete eigenvectors are exactly the discrete
  sines, psi_k(i) = sqrt(2/L) sin(k pi i / n) (sum psi^2 h = 1);
* a dense symmetric eigensolve of the same discrete problem gives the
  eigenvectors of a Morse well (normalised, sign fixed positive next to the
  left wall);
* central finite differences of the oracle's own refined levels give
  dE/dV_i, which must equal psi_i^2 h (Hellmann-Feynman for the discrete
  symmetric problem), also on a trimmed grid;
* the adjoint is linear in the level weights and vanishes outside the
  trimmed window.

A wrong matching factor, a lost rescale, a missing h or a transposed
left/right half fails one of these.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg

import oracle
from tests.helpers import morse_problem
from tests.test_oracle import box_levels, box_problem, discrete_hamiltonian


def test_box_eigenfunctions_are_discrete_sines():
    n, mu, length = 2001, 1.3, 2.5
    prob = box_problem(n, mu, length, e_top=4000.0, m=400)
    assert prob.match == 1  # constant V: minimum at the wall, clamped
    ref = box_levels(n, prob.h, mu, 12)
    _, _, _, lv, _ = oracle.solve(prob, 0, 12, 1e-13, 300)
    i = np.arange(n)
    for k in range(12):
        psi = oracle.eigenfunction(prob, lv[k])
        exact = np.sqrt(2.0 / length) * np.sin((k + 1) * np.pi * i / (n - 1))
        np.testing.assert_allclose(psi, exact, atol=1e-9 * np.abs(exact).max())
        np.testing.assert_allclose(np.sum(psi**2) * prob.h, 1.0, rtol=1e-12)
        np.testing.assert_allclose(lv[k], ref[k], rtol=1e-12)


@pytest.fixture(scope="module")
def dense_morse():
    prob = morse_problem(mu=0.50391, de=38297.0, a=1.9426, re=0.7414, r_min=0.35, r_max=2.5, n=700, m=300)
    eig, vec = scipy.linalg.eigh(discrete_hamiltonian(prob.v, prob.h, prob.mu))
    nb = int(np.sum(eig < prob.e_top))
    _, _, _, lv, _ = oracle.solve(prob, 0, nb, 1e-13, 300)
    return prob, eig, vec, lv, nb


def test_eigenfunctions_match_dense_eigenvectors(dense_morse):
    prob, eig, vec, lv, nb = dense_morse
    assert 1 <= prob.match <= prob.n - 4
    for v in range(nb):
        psi = oracle.eigenfunction(prob, lv[v])
        y = vec[:, v] / np.sqrt(prob.h)
        y = y * np.sign(y[0])
        np.testing.assert_allclose(psi[1:-1], y, atol=2e-9 * np.abs(y).max(), err_msg=f"level {v}")
        assert psi[0] == 0.0 and psi[-1] == 0.0
        assert psi[1] > 0.0
        np.testing.assert_allclose(np.sum(psi**2) * prob.h, 1.0, rtol=1e-12)


def test_hellmann_feynman_against_finite_differences(dense_morse):
    prob, eig, vec, lv, nb = dense_morse
    rng = np.random.default_rng(11)
    levels = [0, 3, nb // 2, nb - 2]
    for v in levels:
        psi = oracle.eigenfunction(prob, lv[v])
        # Points where the level has weight (the derivative is psi^2 h).
        pts = rng.choice(np.flatnonzero(psi**2 * prob.h > 1e-4), 3, replace=False)
        for i in pts:
            d = 1.0
            vals = []
            for s in (+1, -1):
                pv = prob.v.copy()
                pv[i] += s * d
                p2 = oracle.Problem(v=pv, r_min=prob.r_min, r_max=prob.r_max, mu=prob.mu,
                                    e_top=prob.e_top, m=prob.m)
                vals.append(oracle.solve(p2, v, 1, 1e-14, 400)[3][0])
            fd = (vals[0] - vals[1]) / (2 * d)
            np.testing.assert_allclose(psi[i] ** 2 * prob.h, fd, rtol=1e-6, err_msg=f"v={v} i={i}")


def test_adjoint_on_trimmed_grid_and_linearity():
    # A tall inner wall: the grid is trimmed, so the adjoint must vanish on
    # the cut-off points and at both Dirichlet points.
    prob = morse_problem(mu=7.0, de=20000.0, a=2.0, re=1.2, r_min=0.1, r_max=4.0, n=800, m=600)
    first, last, ok = prob.trimmed
    assert ok and first > 0
    _, _, _, lv, _ = oracle.solve(prob, 0, 6, 1e-13, 300)
    g = np.array([1.0, -0.5, 2.0, 0.0, 0.25, 3.0])
    gv = oracle.level_vjp(prob, lv, g)
    acc = np.zeros(prob.n)
    for l in range(6):
        one = np.zeros(6)
        one[l] = 1.0
        col = oracle.level_vjp(prob, lv, one)
        np.testing.assert_allclose(col, oracle.eigenfunction(prob, lv[l]) ** 2 * prob.h, rtol=1e-13, atol=0)
        acc += g[l] * col
    np.testing.assert_allclose(gv, acc, rtol=1e-12, atol=1e-15)
    assert np.all(gv[: first + 1] == 0.0) and np.all(gv[last:] == 0.0)
    # Sum rule: sum_i dE_l/dV_i = 1 (a constant shift of V shifts every level).
    for l in range(6):
        one = np.zeros(6)
        one[l] = 1.0
        np.testing.assert_allclose(oracle.level_vjp(prob, lv, one).sum(), 1.0, rtol=1e-12)
    # Finite differences on the trimmed grid, at the two levels' largest weights.
    for l in (0, 5):
        psi = oracle.eigenfunction(prob, lv[l])
        i = int(np.argmax(psi**2))
        vals = []
        for s in (+1, -1):
            pv = prob.v.copy()
            pv[i] += s * 2.0
            p2 = oracle.Problem(v=pv, r_min=prob.r_min, r_max=prob.r_max, mu=prob.mu, e_top=prob.e_top, m=prob.m)
            assert p2.trimmed == prob.trimmed
            vals.append(oracle.solve(p2, l, 1, 1e-14, 400)[3][0])
        np.testing.assert_allclose(psi[i] ** 2 * prob.h, (vals[0] - vals[1]) / 4.0, rtol=1e-6)


def test_unbound_levels_contribute_nothing():
    prob = morse_problem(mu=0.50391, de=38297.0, a=1.9426, re=0.7414, r_min=0.35, r_max=2.5, n=700, m=300)
    _, _, _, lv, _ = oracle.solve(prob, 0, 40, 1e-12, 300)
    assert np.isnan(lv[-1])
    g = np.ones(40)
    gv = oracle.level_vjp(prob, lv, g)
    ok = np.isfinite(lv)
    np.testing.assert_allclose(gv.sum(), ok.sum(), rtol=1e-12)


def test_parameter_derivatives_match_analytic_morse():
    """sum_i psi_i^2 h dV_i/dtheta equals the analytic Morse dE_v/dtheta
    (dE/dD_e = a sqrt(K'/D_e)(v+1/2), dE/da = 2 sqrt(K' D_e)(v+1/2) - 2 a K'(v+1/2)^2,
    dE/dr_e = 0) when the grid resolves the well and the walls are opaque."""
    from epseon_backend_amd.workloads import MorseSpec

    for s in (MorseSpec(mu=42.342, de=23000.0, a=1.6, re=2.2, r_min=1.6, r_max=5.95, n=50_000, m=2000),
              MorseSpec(mu=7.8, de=31000.0, a=1.9, re=1.1, r_min=0.5, r_max=4.26, n=50_000, m=2000)):
        prob = oracle.Problem(v=s.table(), r_min=s.r_min, r_max=s.r_max, mu=s.mu, e_top=s.de, m=s.m)
        _, _, _, lv, _ = oracle.solve(prob, 0, 3, 1e-13, 300)
        r = s.r_min + np.arange(s.n) * prob.h
        ex = np.exp(-s.a * (r - s.re))
        d_de, d_a, d_re = (1 - ex) ** 2, 2 * s.de * (1 - ex) * ex * (r - s.re), -2 * s.de * s.a * (1 - ex) * ex
        kp = oracle.K_CM / s.mu
        for v in range(3):
            w = oracle.eigenfunction(prob, lv[v]) ** 2 * prob.h
            x = v + 0.5
            a_de = s.a * np.sqrt(kp / s.de) * x
            a_a = 2 * np.sqrt(kp * s.de) * x - 2 * s.a * kp * x * x
            np.testing.assert_allclose(w @ d_de, a_de, rtol=1e-10)
            np.testing.assert_allclose(w @ d_a, a_a, rtol=1e-10)
            assert abs(w @ d_re) < 1e-9 * abs(a_a) * s.a
