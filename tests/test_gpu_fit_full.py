"""The differentiation workload at full size on the GPU.

256 Morse potentials drawn as in the batched workload (seed + 1) on
50,000-point FP64 grids, levels v = 0..40, loss = sum of squared residuals
against levels of +-2 % perturbed parameters plus N(0, 0.01^2) noise.  The
backward pass (adjoint kernel, then torch autograd to D_e, a, r_e) over the
whole batch must be finite and deterministic (two passes bit-identical); the
parameter gradient of sampled potentials must match central differences of
whole-batch solves (step 1e-6 relative); the pointwise gradient dL/dV of
sampled potentials must match the oracle's adjoint at the GPU's own levels.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from epseon_backend_amd.workloads import DIFF_SEED, differentiation, perturbed

pytestmark = pytest.mark.gpu


def _log(msg: str) -> None:
    sys.stdout.write(msg + "\n")
    sys.stdout.flush()


def test_fit_workload_full_batch():
    import torch

    from epseon_backend_amd import capi
    from epseon_backend_amd.autograd import LevelSolver, LevelSpec, morse_table, numerov_levels, uniform_grid

    if capi.device_count() < 1:
        pytest.fail("no HIP device visible")
    _log("case: fitting workload, 256 potentials x 50000 points, levels 0..40, rel_tol 1e-14")
    levels = 41
    specs = differentiation(256)
    obs = perturbed(specs, seed=DIFF_SEED + 1)
    lspecs = [LevelSpec(s.r_min, s.r_max, s.mu, s.de, s.m, 0, levels) for s in specs]
    dev = torch.device("cuda:0")
    t = lambda x: torch.tensor(x, dtype=torch.float64, device=dev)  # noqa: E731
    r = uniform_grid(t([s.r_min for s in specs]), t([s.r_max for s in specs]), specs[0].n)
    solver = LevelSolver(0, rel_tol=1e-14, max_iter=300)
    with torch.no_grad():
        e_obs = numerov_levels(morse_table(r, t([s.de for s in obs]), t([s.a for s in obs]),
                                           t([s.re for s in obs])), lspecs, solver)
        gen = torch.Generator(device=dev).manual_seed(DIFF_SEED)
        e_obs = e_obs + 0.01 * torch.randn(e_obs.shape, dtype=torch.float64, device=dev, generator=gen)
    theta0 = [t([s.de for s in specs]), t([s.a for s in specs]), t([s.re for s in specs])]

    def loss_of(de, a, re):
        e = numerov_levels(morse_table(r, de, a, re), lspecs, solver)
        ok = torch.isfinite(e) & torch.isfinite(e_obs)
        return torch.where(ok, (e - e_obs) ** 2, torch.zeros_like(e)).sum(), e

    grads = []
    for _ in range(2):
        th = [x.clone().requires_grad_() for x in theta0]
        loss, e = loss_of(*th)
        loss.backward()
        grads.append([x.grad.detach().cpu().numpy().copy() for x in th])
    n_levels = int(torch.isfinite(e).sum())
    _log(f"  loss {float(loss.detach()):.6g}, {n_levels} converged levels")
    assert n_levels > 8000
    for a, b in zip(grads[0], grads[1]):
        assert np.all(np.isfinite(a))
        np.testing.assert_array_equal(a.view(np.int64), b.view(np.int64))
    g = grads[0]

    # Central differences of whole-batch solves for sampled potentials: the
    # other potentials' levels are bit-identical in both solves.
    rng = np.random.default_rng(3)
    usable = [j for j in range(len(specs)) if np.isfinite(e[j].detach().cpu().numpy()).any()]
    for j in sorted(rng.choice(usable, 6, replace=False)):
        scale = max(abs(g[q][j] * float(theta0[q][j])) for q in range(3))
        for k in range(3):
            vals, pos = [], []
            for sgn in (+1, -1):
                args = [x.clone() for x in theta0]
                args[k][j] = args[k][j] * (1.0 + sgn * 1e-6)
                pos.append(float(args[k][j]))
                with torch.no_grad():
                    vals.append(float(loss_of(*args)[0]))
            fd = (vals[0] - vals[1]) / (pos[0] - pos[1])
            assert abs(g[k][j] - fd) * float(theta0[k][j]) <= 1e-6 * scale, (j, k, g[k][j], fd)

    # Pointwise adjoint of sampled potentials against the oracle.
    v = morse_table(r, *theta0).detach().requires_grad_()
    ev = numerov_levels(v, lspecs, solver)
    ok = torch.isfinite(ev) & torch.isfinite(e_obs)
    loss_v = torch.where(ok, (ev - e_obs) ** 2, torch.zeros_like(ev)).sum()
    loss_v.backward()
    gv = v.grad.cpu().numpy()
    lv = ev.detach().cpu().numpy()
    dl = torch.where(ok, 2 * (ev - e_obs), torch.zeros_like(ev)).detach().cpu().numpy()
    table = v.detach().cpu().numpy()
    for j in sorted(rng.choice(usable, 4, replace=False)):
        s = specs[j]
        prob = oracle.Problem(v=table[j], r_min=s.r_min, r_max=s.r_max, mu=s.mu, e_top=s.de, m=s.m)
        ref = oracle.level_vjp(prob, lv[j], dl[j])
        np.testing.assert_allclose(gv[j], ref, rtol=0, atol=1e-11 * np.abs(ref).max(), err_msg=f"potential {j}")
    solver.close()
