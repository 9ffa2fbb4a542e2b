"""Eigenfunctions and the adjoint on the GPU, through the C ABI and the
torch.autograd front end.

* eigenfunctions (epseon_get_wavefunctions) against the oracle's eo_psi at
  the GPU's own levels (which are bit-identical to the oracle's), and their
  orthonormality;
* the adjoint (epseon_level_vjp) against the oracle's eo_level_vjp, host
  and device-pointer forms;
* Hellmann-Feynman against central finite differences of GPU solves, both
  pointwise in V and, through torch autograd on a Morse tabulation, in
  (D_e, a, r_e) with the 1e-6 relative step of the fitting workload;
* per-level parameter derivatives against the analytic Morse ones.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from epseon_backend_amd.workloads import H2, differentiation, perturbed
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


def _solve(capi, specs, counts, rel_tol=1e-10, mode=None):
    ctx = capi.Context(0)
    if mode is not None:
        ctx.set_refine_mode(mode)
    ctx.load([capi.Potential(values=s.table(), r_min=s.r_min, r_max=s.r_max, reduced_mass=s.mu, e_top=s.de,
                             n_energies=s.m, level_min=0, level_count=c) for s, c in zip(specs, counts)])
    ctx.solve(rel_tol, 300)
    return ctx


def _batch_specs():
    specs = differentiation(64, n_points=6000)[:12]
    return [s for s in specs if problem_from_spec(s).usable]


def test_eigenfunctions_match_oracle(capi):
    _log("case: eigenfunctions, H2 (10000 points, 17 levels) + 12 fitting-batch potentials at 6000 points")
    specs = [H2] + _batch_specs()
    counts = [17] + [min(41, s.bound_levels()) for s in specs[1:]]
    ctx = _solve(capi, specs, counts)
    lv, st, _ = ctx.levels()
    lo = 0
    for p, s in enumerate(specs):
        psi = ctx.wavefunctions(p)
        assert psi.shape == (counts[p], s.n)
        prob = problem_from_spec(s)
        for l in range(counts[p]):
            if st[lo + l] != 0:
                assert np.all(psi[l] == 0.0)
                continue
            ref = oracle.eigenfunction(prob, lv[lo + l])
            np.testing.assert_allclose(psi[l], ref, rtol=0, atol=1e-12 * np.abs(ref).max(),
                                       err_msg=f"potential {p} level {l}")
            assert psi[l][prob.trimmed[0] + 1] > 0.0
        lo += counts[p]
    # Orthonormality (the psi_v are eigenvectors of one symmetric matrix),
    # independent of the oracle: levels at 1e-10 relative put the Gram matrix
    # within ~1e-9 of the identity.
    for p in (0, 1):
        psi = ctx.wavefunctions(p)
        h = (specs[p].r_max - specs[p].r_min) / (specs[p].n - 1)
        gram = psi @ psi.T * h
        err = np.abs(gram - np.eye(len(psi))).max()
        _log(f"  potential {p}: max |Gram - I| = {err:.2e} over {len(psi)} levels")
        assert err < 1e-8
    stats = ctx.stats()
    _log(f"  normalize {stats.normalize_ms:.3f} ms over {stats.normalize_point_updates:.3g} point-updates")
    # One launch for every potential: the same bits, in load order.
    everything = ctx.all_wavefunctions()
    for p in range(len(specs)):
        np.testing.assert_array_equal(everything[p], ctx.wavefunctions(p))
    ctx.close()


def test_vjp_matches_oracle_host_and_device(capi):
    import torch

    _log("case: level adjoint vs oracle, same batch, random level weights")
    specs = [H2] + _batch_specs()
    counts = [17] + [min(41, s.bound_levels()) for s in specs[1:]]
    ctx = _solve(capi, specs, counts)
    lv, st, _ = ctx.levels()
    rng = np.random.default_rng(3)
    g = rng.normal(size=lv.shape)
    gv = ctx.level_vjp(g)
    # Device pointers: same result into a torch buffer.
    gd = torch.from_numpy(g).cuda()
    out = torch.full((ctx.v_span,), float("nan"), dtype=torch.float64, device="cuda")
    ctx.level_vjp_device(gd.data_ptr(), out.data_ptr())
    np.testing.assert_array_equal(out.cpu().numpy(), gv)
    lo = 0
    for p, s in enumerate(specs):
        prob = problem_from_spec(s)
        lev = np.where(st[lo : lo + counts[p]] == 0, lv[lo : lo + counts[p]], np.nan)
        ref = oracle.level_vjp(prob, lev, g[lo : lo + counts[p]])
        seg = gv[ctx.v_offsets[p] : ctx.v_offsets[p] + s.n]
        np.testing.assert_allclose(seg, ref, rtol=0, atol=1e-11 * np.abs(ref).max(), err_msg=f"potential {p}")
        first, last, _ = prob.trimmed
        assert np.all(seg[: first + 1] == 0.0) and np.all(seg[last:] == 0.0)
        lo += counts[p]
    stats = ctx.stats()
    _log(f"  vjp {stats.vjp_ms:.3f} ms")
    ctx.close()


def test_vjp_covers_every_point(capi):
    """Device-pointer adjoint into a NaN-filled buffer: potentials without
    levels and unresolved potentials get zeros everywhere, the others match
    the oracle; no point is left unwritten."""
    import torch

    from epseon_backend_amd.workloads import MorseSpec

    _log("case: adjoint over a batch with a zero-level and an unresolved potential, NaN-filled output")
    specs = [H2, MorseSpec(2.0, 8000.0, 1.2, 1.2, 0.65, 7.0, 900, 64),
             MorseSpec(50.0, 40000.0, 2.5, 2.0, 0.2, 9.0, 60, 64)] + _batch_specs()[:4]
    counts = [17, 0, 3] + [min(41, s.bound_levels()) for s in specs[3:]]
    ctx = _solve(capi, specs, counts)
    assert ctx.potential_info(2).flags == 1
    lv, st, _ = ctx.levels()
    g = np.random.default_rng(11).normal(size=lv.shape)
    gd = torch.from_numpy(g).cuda()
    out = torch.full((ctx.v_span,), float("nan"), dtype=torch.float64, device="cuda")
    ctx.level_vjp_device(gd.data_ptr(), out.data_ptr())
    gv = out.cpu().numpy()
    assert np.all(np.isfinite(gv))
    lo = 0
    for p, s in enumerate(specs):
        seg = gv[ctx.v_offsets[p] : ctx.v_offsets[p] + s.n]
        if p in (1, 2):
            assert np.all(seg == 0.0), p
        else:
            prob = problem_from_spec(s)
            lev = np.where(st[lo : lo + counts[p]] == 0, lv[lo : lo + counts[p]], np.nan)
            ref = oracle.level_vjp(prob, lev, g[lo : lo + counts[p]])
            np.testing.assert_allclose(seg, ref, rtol=0, atol=1e-11 * np.abs(ref).max(), err_msg=f"potential {p}")
        lo += counts[p]
    ctx.close()


def test_pointwise_hellmann_feynman_vs_finite_differences(capi):
    _log("case: dE/dV_i vs central differences of GPU solves, H2, levels 0, 5, 12")
    ctx = _solve(capi, [H2], [17], rel_tol=1e-14)
    lv, _, _ = ctx.levels()
    psi = ctx.wavefunctions(0)
    h = (H2.r_max - H2.r_min) / (H2.n - 1)
    base = H2.table()
    for v in (0, 5, 12):
        i = int(np.argmax(psi[v] ** 2))
        vals = []
        for sgn in (+1, -1):
            pv = base.copy()
            pv[i] += sgn * 1.0
            c2 = capi.Context(0)
            c2.load([capi.Potential(values=pv, r_min=H2.r_min, r_max=H2.r_max, reduced_mass=H2.mu, e_top=H2.de,
                                    n_energies=H2.m, level_min=v, level_count=1)])
            c2.solve(1e-14, 300)
            vals.append(c2.levels()[0][0])
            c2.close()
        fd = (vals[0] - vals[1]) / 2.0
        np.testing.assert_allclose(psi[v][i] ** 2 * h, fd, rtol=1e-6, err_msg=f"level {v} point {i}")
    ctx.close()


def _theta_setup(torch, specs):
    p = len(specs)
    dev = torch.device("cuda")
    t = lambda x: torch.tensor(x, dtype=torch.float64, device=dev)  # noqa: E731
    de = t([s.de for s in specs])
    a = t([s.a for s in specs])
    re = t([s.re for s in specs])
    r_min = t([s.r_min for s in specs])
    r_max = t([s.r_max for s in specs])
    return p, de, a, re, r_min, r_max


def test_parameter_gradient_vs_finite_differences(capi):
    """Loss of the fitting workload, gradient in (D_e, a, r_e) through the
    adjoint and torch autograd, against central differences with step 1e-6
    relative (GPU solves at rel_tol 1e-14)."""
    import torch

    from epseon_backend_amd.autograd import LevelSolver, LevelSpec, morse_table, numerov_levels, uniform_grid

    _log("case: theta-gradient of the fitting loss vs central differences, 4 potentials x 50000 points")
    specs = [s for s in differentiation(16) if problem_from_spec(s).usable][:4]
    obs_specs = perturbed(specs, seed=7)
    levels = 6
    lspecs = [LevelSpec(s.r_min, s.r_max, s.mu, s.de, s.m, 0, levels) for s in specs]
    solver = LevelSolver(0, rel_tol=1e-14, max_iter=300)
    p, de, a, re, r_min, r_max = _theta_setup(torch, specs)
    r = uniform_grid(r_min, r_max, specs[0].n)
    _, de_o, a_o, re_o, _, _ = _theta_setup(torch, obs_specs)
    with torch.no_grad():
        e_obs = numerov_levels(morse_table(r, de_o, a_o, re_o), lspecs, solver)
        e_obs = e_obs + 0.01 * torch.randn(e_obs.shape, dtype=torch.float64, device="cuda",
                                          generator=torch.Generator(device="cuda").manual_seed(5))

    def loss_of(de_, a_, re_):
        e = numerov_levels(morse_table(r, de_, a_, re_), lspecs, solver)
        ok = torch.isfinite(e) & torch.isfinite(e_obs)
        return torch.where(ok, (e - e_obs) ** 2, torch.zeros_like(e)).sum()

    th = [de.clone().requires_grad_(), a.clone().requires_grad_(), re.clone().requires_grad_()]
    loss = loss_of(*th)
    loss.backward()
    grads = [x.grad.cpu().numpy() for x in th]
    for k in range(3):
        for j in range(p):
            vals = []
            for sgn in (+1, -1):
                args = [x.detach().clone() for x in th]
                args[k][j] = args[k][j] * (1.0 + sgn * 1e-6)
                with torch.no_grad():
                    vals.append(float(loss_of(*args)))
            step = float(th[k][j].detach()) * 1e-6
            fd = (vals[0] - vals[1]) / (2 * step)
            # Compare sensitivities to a relative change of each parameter.
            scale = max(abs(grads[q][j] * float(th[q][j].detach())) for q in range(3))
            assert abs(grads[k][j] - fd) * float(th[k][j].detach()) <= 1e-6 * scale, (k, j, grads[k][j], fd)
    _log(f"  loss {float(loss.detach()):.6g}; adjoint pass {solver.last_stats.vjp_ms:.3f} ms")
    solver.close()


def test_level_derivatives_match_analytic_morse(capi):
    """Per-level dE_v/dtheta from one-hot backward passes against the analytic
    Morse derivatives, on potentials whose walls are opaque."""
    import torch

    from epseon_backend_amd.autograd import LevelSolver, LevelSpec, morse_table, numerov_levels, uniform_grid

    _log("case: per-level theta derivatives vs analytic Morse, fitting-batch potentials at 50000 points")
    cand = [s for s in differentiation(64) if problem_from_spec(s).usable and s.mu > 2.0][:6]
    levels = 3
    lspecs = [LevelSpec(s.r_min, s.r_max, s.mu, s.de, s.m, 0, levels) for s in cand]
    solver = LevelSolver(0, rel_tol=1e-13)
    p, de, a, re, r_min, r_max = _theta_setup(torch, cand)
    r = uniform_grid(r_min, r_max, cand[0].n)
    th = [de.clone().requires_grad_(), a.clone().requires_grad_(), re.clone().requires_grad_()]
    e = numerov_levels(morse_table(r, *th), lspecs, solver)
    mu = torch.tensor([s.mu for s in cand], dtype=torch.float64, device="cuda")
    kp = oracle.K_CM / mu
    for v in range(levels):
        g = torch.zeros_like(e)
        g[:, v] = 1.0
        d_de, d_a, d_re = torch.autograd.grad(e, th, g, retain_graph=True)
        x = v + 0.5
        a_de = th[1] * torch.sqrt(kp / th[0]) * x
        a_a = 2 * torch.sqrt(kp * th[0]) * x - 2 * th[1] * kp * x * x
        np.testing.assert_allclose(d_de.cpu().numpy(), a_de.detach().cpu().numpy(), rtol=1e-8)
        np.testing.assert_allclose(d_a.cpu().numpy(), a_a.detach().cpu().numpy(), rtol=1e-8)
        assert np.all(np.abs(d_re.cpu().numpy()) < 1e-7 * np.abs(a_a.detach().cpu().numpy()))
    solver.close()
