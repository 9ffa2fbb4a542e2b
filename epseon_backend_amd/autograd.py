"""torch.autograd front end of the Numerov levels and their adjoint.

``numerov_levels(v, specs)`` solves the bound levels of a batch of tabulated
potentials on the GPU (scan + bracket + refine through libepseon_hip) and
returns them as a differentiable tensor.  Its backward is the adjoint kernel
``epseon_level_vjp``: for every converged level the eigenfunction is
recomputed on the device (two inward shots matched at min V) and

    dL/dV_i = sum_v dL/dE_v * psi_v(i)^2 * h,

the Hellmann-Feynman derivative of the discrete Numerov eigenproblem (exact
for the discretised problem).  Nothing of psi is stored between forward and
backward.

Chaining to potential parameters is ordinary torch autograd: tabulate V(theta)
with torch ops (``morse_table`` does D_e (1 - exp(-a (r - r_e)))^2) on a grid
that does not depend on theta, and backpropagate through both.

The reference has no gradient path; its Python surface
(python/epseon_backend/device/gpu/_libepseon_gpu.pyi) and the algorithm record
(cpp/gpu/include/epseon/gpu/task_configurator/algorithm_config.hpp:75-199)
define the inputs this module takes: grid (min_r, max_r, point_count),
reduced mass from the two atom masses, the energy window below the asymptote
and the level range.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Sequence

import numpy as np
import torch

from . import capi


@dataclass
class LevelSpec:
    """Solver settings of one potential (grid, mass, window, levels)."""

    r_min: float
    r_max: float
    reduced_mass: float
    e_top: float
    n_energies: int
    level_min: int = 0
    level_count: int = 1


class LevelSolver:
    """A device context reused across calls (its arena and stream persist).

    Each forward loads the batch again; the backward of the most recent
    forward runs on the loaded state, an older one re-solves its saved V
    first (results are deterministic, so this only costs time)."""

    def __init__(self, device: int | torch.device | None = None, rel_tol: float = 1e-10,
                 max_iter: int = 200, refine_mode: int = capi.REFINE_AUTO) -> None:
        if device is None:
            device = torch.cuda.current_device()
        if isinstance(device, torch.device):
            device = device.index if device.index is not None else torch.cuda.current_device()
        self.device = int(device)
        self.rel_tol = float(rel_tol)
        self.max_iter = int(max_iter)
        self.ctx = capi.Context(self.device)
        self.ctx.set_refine_mode(refine_mode)
        self.generation = 0
        self.last_stats: capi.Stats | None = None

    def close(self) -> None:
        self.ctx.close()

    def _load_and_solve(self, v: torch.Tensor, specs: Sequence[LevelSpec]) -> None:
        """Upload the rows of v: straight from device memory when v lives on
        this solver's GPU, through the host otherwise."""
        rows, n = v.shape
        meta = [
            capi.Potential(values=range(n), r_min=s.r_min, r_max=s.r_max, reduced_mass=s.reduced_mass,
                           e_top=s.e_top, n_energies=s.n_energies, level_min=s.level_min,
                           level_count=s.level_count)
            for s in specs
        ]
        if v.is_cuda and v.device.index == self.device:
            vd = v.detach().to(torch.float64).contiguous()
            torch.cuda.current_stream(vd.device).synchronize()
            self.ctx.load_device(meta, vd.data_ptr(), [i * n for i in range(rows)])
        else:
            v_host = v.detach().to("cpu", torch.float64).contiguous().numpy()
            for i, m in enumerate(meta):
                m.values = v_host[i]
            self.ctx.load(meta)
        self.ctx.solve(self.rel_tol, self.max_iter)

    def solve(self, v: torch.Tensor, specs: Sequence[LevelSpec]) -> tuple[torch.Tensor, int]:
        """Levels of the rows of v (P, N) as a (P, L) float64 tensor on v's
        device, NaN where a level did not converge; plus the generation tag."""
        self._load_and_solve(v, specs)
        lv, st, _ = self.ctx.levels()
        lv = np.where(st == capi.LEVEL_OK, lv, np.nan)
        self.generation += 1
        counts = {s.level_count for s in specs}
        if len(counts) != 1:
            raise ValueError("numerov_levels needs the same level_count for every potential")
        out = torch.from_numpy(lv.reshape(len(specs), counts.pop())).to(v.device)
        return out, self.generation

    def vjp(self, v: torch.Tensor, specs: Sequence[LevelSpec], generation: int,
            grad_levels: torch.Tensor) -> torch.Tensor:
        """sum_l grad_levels[l] dE_l/dV as a tensor shaped like v."""
        if generation != self.generation:
            self._load_and_solve(v, specs)
            self.generation += 1
        g = torch.nan_to_num(grad_levels.detach().to(torch.float64), nan=0.0, posinf=0.0, neginf=0.0)
        if v.is_cuda:
            g = g.to(v.device).contiguous()
            out = torch.empty(v.shape[0] * v.shape[1], dtype=torch.float64, device=v.device)
            torch.cuda.current_stream(v.device).synchronize()
            self.ctx.level_vjp_device(g.data_ptr(), out.data_ptr())
        else:
            out = torch.from_numpy(self.ctx.level_vjp(g.reshape(-1).cpu().numpy()))
        self.last_stats = self.ctx.stats()
        return out.reshape(v.shape)


class _NumerovLevels(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v: torch.Tensor, solver: LevelSolver, specs: tuple[LevelSpec, ...]) -> torch.Tensor:
        levels, gen = solver.solve(v, specs)
        ctx.solver = solver
        ctx.specs = specs
        ctx.generation = gen
        ctx.save_for_backward(v)
        return levels

    @staticmethod
    def backward(ctx, grad_levels: torch.Tensor):
        (v,) = ctx.saved_tensors
        grad_v = ctx.solver.vjp(v, ctx.specs, ctx.generation, grad_levels.reshape(-1))
        return grad_v.to(v.dtype), None, None


def numerov_levels(v: torch.Tensor, specs: Sequence[LevelSpec] | LevelSpec,
                   solver: LevelSolver | None = None) -> torch.Tensor:
    """Bound levels (P, L) of the potentials v (P, N) (cm^-1 on uniform grids),
    differentiable with respect to v through the adjoint kernel."""
    if v.dim() != 2:
        raise ValueError("v must be a (potentials, points) tensor")
    if isinstance(specs, LevelSpec):
        specs = [specs] * v.shape[0]
    if len(specs) != v.shape[0]:
        raise ValueError("one LevelSpec per potential is needed")
    if solver is None:
        solver = LevelSolver(v.device if v.is_cuda else None)
    return _NumerovLevels.apply(v, solver, tuple(specs))


def uniform_grid(r_min: torch.Tensor, r_max: torch.Tensor, n: int) -> torch.Tensor:
    """r_i = r_min + i h, h = (r_max - r_min)/(n - 1), for each potential (P, n)."""
    i = torch.arange(n, dtype=torch.float64, device=r_min.device)
    h = (r_max - r_min) / (n - 1)
    return r_min[:, None] + i[None, :] * h[:, None]


def morse_table(r: torch.Tensor, de: torch.Tensor, a: torch.Tensor, re: torch.Tensor) -> torch.Tensor:
    """V = D_e (1 - exp(-a (r - r_e)))^2 per row of r (parameters are (P,))."""
    x = 1.0 - torch.exp(-a[:, None] * (r - re[:, None]))
    return de[:, None] * x * x
