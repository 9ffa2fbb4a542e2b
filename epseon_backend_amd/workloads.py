"""Synthetic workloads of the Numerov path (shared by bench.py and the tests).

Potentials are Morse curves V(r) = D_e (1 - exp(-a (r - r_e)))^2, optionally
plus a Gaussian barrier B exp(-((r - r_b)/w)^2) (double wells), tabulated in
FP64 with numpy on the uniform grid r_i = r_min + i h.  Units: cm^-1, A, amu.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

K_CM = 16.857629191640175  # hbar^2/(2 amu A^2) in cm^-1


@dataclass
class MorseSpec:
    mu: float
    de: float
    a: float
    re: float
    r_min: float
    r_max: float
    n: int
    m: int
    barrier: float = 0.0  # B (cm^-1); 0 = plain Morse
    r_b: float = 0.0
    width: float = 1.0

    def grid(self) -> np.ndarray:
        h = (self.r_max - self.r_min) / (self.n - 1)
        return self.r_min + np.arange(self.n, dtype=np.float64) * h

    def table(self) -> np.ndarray:
        r = self.grid()
        x = 1.0 - np.exp(-self.a * (r - self.re))
        v = self.de * (x * x)
        if self.barrier != 0.0:
            z = (r - self.r_b) / self.width
            v = v + self.barrier * np.exp(-(z * z))
        return v

    def bound_levels(self) -> int:
        """Number of analytic Morse levels below D_e (infinite domain)."""
        lam = np.sqrt(self.de * self.mu / K_CM) / self.a
        return int(np.floor(lam - 0.5)) + 1

    def analytic(self, v: np.ndarray) -> np.ndarray:
        kp = K_CM / self.mu
        x = np.asarray(v, dtype=np.float64) + 0.5
        return 2.0 * self.a * np.sqrt(kp * self.de) * x - self.a * self.a * kp * x * x


# The light single-potential case (H2-like) and the heavy one (I2-like).
H2 = MorseSpec(mu=0.50391, de=38297.0, a=1.9426, re=0.7414, r_min=0.2, r_max=5.0, n=10_000, m=2048)
I2 = MorseSpec(mu=63.452, de=12440.0, a=1.8676, re=2.666, r_min=2.0, r_max=8.0, n=200_000, m=65_536)

BATCH_SEED = 20261202


def batched(count: int = 4096, seed: int = BATCH_SEED,
            n_range: tuple[int, int] = (2000, 64000), m_range: tuple[int, int] = (256, 8192)) -> list[MorseSpec]:
    """Seeded batch: mu log-uniform [0.5, 100] amu, D_e U[5e3, 5e4] cm^-1,
    a U[1, 2.5] 1/A, r_e U[0.7, 3] A, N and M log-uniform, grid
    [r_e - 0.6, r_e + 6/a].  Draw order: mu, D_e, a, r_e, N, M (each a vector
    of `count` values from numpy.random.default_rng(seed))."""
    rng = np.random.default_rng(seed)
    mu = np.exp(rng.uniform(np.log(0.5), np.log(100.0), count))
    de = rng.uniform(5e3, 5e4, count)
    a = rng.uniform(1.0, 2.5, count)
    re = rng.uniform(0.7, 3.0, count)
    n = np.floor(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), count))).astype(np.int64)
    m = np.floor(np.exp(rng.uniform(np.log(m_range[0]), np.log(m_range[1]), count))).astype(np.int64)
    return [
        MorseSpec(float(mu[i]), float(de[i]), float(a[i]), float(re[i]), float(re[i] - 0.6),
                  float(re[i] + 6.0 / a[i]), int(n[i]), int(m[i]))
        for i in range(count)
    ]


DIFF_SEED = BATCH_SEED + 1


def differentiation(count: int = 256, seed: int = DIFF_SEED, n_points: int = 50_000) -> list[MorseSpec]:
    """Fitting batch: potentials drawn as in ``batched`` (seed + 1 by default),
    every grid replaced by ``n_points`` points on the same [r_e - 0.6, r_e + 6/a]."""
    specs = batched(count, seed)
    for s in specs:
        s.n = n_points
    return specs


def perturbed(specs: list[MorseSpec], seed: int, rel: float = 0.02) -> list[MorseSpec]:
    """Copies with D_e, a, r_e each scaled by (1 + rel U[-1, 1]) (draw order
    D_e, a, r_e, each a vector over the batch); grids, masses and counts kept."""
    rng = np.random.default_rng(seed)
    k = len(specs)
    f_de = 1.0 + rel * rng.uniform(-1.0, 1.0, k)
    f_a = 1.0 + rel * rng.uniform(-1.0, 1.0, k)
    f_re = 1.0 + rel * rng.uniform(-1.0, 1.0, k)
    return [
        MorseSpec(s.mu, s.de * f_de[i], s.a * f_a[i], s.re * f_re[i], s.r_min, s.r_max, s.n, s.m)
        for i, s in enumerate(specs)
    ]


STRESS_SEED = BATCH_SEED + 2


def double_wells(count: int = 64, seed: int = STRESS_SEED, n_points: int = 1_000_000,
                 n_energies: int = 262_144) -> list[MorseSpec]:
    """Stress batch: Morse parameters drawn as in ``batched`` (same draw order,
    this seed), plus a Gaussian barrier B exp(-((r - r_b)/w)^2) with
    B ~ U[500, 3000] cm^-1, r_b = r_e + U[0.3, 0.8] A, w ~ U[0.05, 0.2] A drawn
    afterwards from the same generator (order B, r_b - r_e, w, each a vector);
    every grid has ``n_points`` points on [r_e - 0.6, r_e + 6/a] and
    ``n_energies`` trial energies."""
    rng = np.random.default_rng(seed)
    mu = np.exp(rng.uniform(np.log(0.5), np.log(100.0), count))
    de = rng.uniform(5e3, 5e4, count)
    a = rng.uniform(1.0, 2.5, count)
    re = rng.uniform(0.7, 3.0, count)
    _ = rng.uniform(0.0, 1.0, count)  # grid-length draw of the batched order (unused: fixed N)
    _ = rng.uniform(0.0, 1.0, count)  # trial-count draw of the batched order (unused: fixed M)
    b = rng.uniform(500.0, 3000.0, count)
    off = rng.uniform(0.3, 0.8, count)
    w = rng.uniform(0.05, 0.2, count)
    return [
        MorseSpec(float(mu[i]), float(de[i]), float(a[i]), float(re[i]), float(re[i] - 0.6),
                  float(re[i] + 6.0 / a[i]), n_points, n_energies, float(b[i]), float(re[i] + off[i]),
                  float(w[i]))
        for i in range(count)
    ]
