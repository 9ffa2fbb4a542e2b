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


def problem_from_spec(spec) -> oracle.Problem:
    """oracle.Problem on the numpy-tabulated table of a workloads.MorseSpec."""
    return oracle.Problem(v=spec.table(), r_min=spec.r_min, r_max=spec.r_max, mu=spec.mu,
                          e_top=spec.de, m=spec.m)


def random_morse_batch(count: int, seed: int, n_range=(2000, 64000), m_range=(256, 8192)):
    """Seeded batch drawn as in the batched workload (epseon_backend_amd.workloads.batched)."""
    from epseon_backend_amd.workloads import batched

    return [problem_from_spec(s) for s in batched(count, seed, n_range, m_range)]
