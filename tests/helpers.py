"""Workload builders shared by the tests (potential tables and problems)."""
from __future__ import annotations

import numpy as np

import oracle

# Single Morse potentials quoted by the workload description
# (mu amu, D_e cm^-1, a 1/A, r_e A, r_min, r_max, N, M).
CFG_H2 = dict(mu=0.50391, de=38297.0, a=1.9426, re=0.7414, r_min=0.2, r_max=5.0, n=10_000, m=2048)
CFG_I2 = dict(mu=63.452, de=12440.0, a=1.8676, re=2.666, r_min=2.0, r_max=8.0, n=200_000, m=65_536)


def morse_problem(mu, de, a, re, r_min, r_max, n, m) -> oracle.Problem:
    h = oracle.grid_step(r_min, r_max, n)
    v = oracle.morse(de, re, a, r_min, h, n)
    return oracle.Problem(v=v, r_min=r_min, r_max=r_max, mu=mu, e_top=de, m=m)


def random_morse_batch(count: int, seed: int, n_range=(2000, 64000), m_range=(256, 8192)):
    """Seeded batch drawn as in the batched workload: mu log-uniform [0.5, 100],
    D_e U[5e3, 5e4], a U[1, 2.5], r_e U[0.7, 3], N and M log-uniform; grid
    [r_e - 0.6, r_e + 6/a].  Draw order: mu, D_e, a, r_e, N, M."""
    rng = np.random.default_rng(seed)
    mu = np.exp(rng.uniform(np.log(0.5), np.log(100.0), count))
    de = rng.uniform(5e3, 5e4, count)
    a = rng.uniform(1.0, 2.5, count)
    re = rng.uniform(0.7, 3.0, count)
    n = np.floor(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), count))).astype(np.int64)
    m = np.floor(np.exp(rng.uniform(np.log(m_range[0]), np.log(m_range[1]), count))).astype(np.int64)
    out = []
    for i in range(count):
        out.append(
            morse_problem(mu[i], de[i], a[i], re[i], re[i] - 0.6, re[i] + 6.0 / a[i], int(n[i]), int(m[i]))
        )
    return out
