"""Potential tabulation against 30-digit mpmath values (tests/golden/
potential_tables.json, made by tests/golden/make_golden.py): the oracle's C
Morse table, the workloads' numpy tables (Morse and double well) and the torch
front end's Morse table with its autograd derivatives d V / d(D_e, a, r_e), at
~20 FP64 grid points of the H2-like and I2-like grids and of one double well of
the stress batch.  Also pins the seeded batch generator's draw order."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle
from epseon_backend_amd.autograd import morse_table, uniform_grid
from epseon_backend_amd.workloads import BATCH_SEED, MorseSpec, batched

GOLDEN = Path(__file__).resolve().parent / "golden"
EPS = np.finfo(np.float64).eps


@pytest.fixture(scope="module")
def tables():
    return json.loads((GOLDEN / "potential_tables.json").read_text())


def _spec(t) -> MorseSpec:
    return MorseSpec(mu=1.0, de=float(t["de"]), a=float(t["a"]), re=float(t["re"]), r_min=float(t["r_min"]),
                     r_max=float(t["r_max"]), n=int(t["n"]), m=1, barrier=float(t["barrier"]),
                     r_b=float(t["r_b"]), width=float(t["width"]))


def _golden(t, key):
    idx = np.array([p["i"] for p in t["points"]])
    return idx, np.array([float(p[key]) for p in t["points"]])


def _close(got, want, scale):
    # A few ulp of the table's scale: exp and the (1 - e^-x)^2 cancellation
    # near r_e are accurate to about 1 ulp of D_e.
    np.testing.assert_allclose(got, want, rtol=8 * EPS, atol=16 * EPS * scale)


@pytest.mark.parametrize("name", ["h2", "i2", "double_well"])
def test_grid_points_are_the_golden_ones(tables, name):
    t = tables[name]
    s = _spec(t)
    idx, _ = _golden(t, "v")
    r = s.grid()[idx]
    np.testing.assert_array_equal(r, np.array([float(p["r"]) for p in t["points"]]))


@pytest.mark.parametrize("name", ["h2", "i2", "double_well"])
def test_numpy_tables(tables, name):
    t = tables[name]
    s = _spec(t)
    idx, want = _golden(t, "v")
    _close(s.table()[idx], want, s.de + s.barrier)


@pytest.mark.parametrize("name", ["h2", "i2"])
def test_oracle_morse_table(tables, name):
    t = tables[name]
    s = _spec(t)
    idx, want = _golden(t, "v")
    h = oracle.grid_step(s.r_min, s.r_max, s.n)
    v = oracle.morse(s.de, s.re, s.a, s.r_min, h, s.n)
    _close(v[idx], want, s.de)


@pytest.mark.parametrize("name", ["h2", "i2"])
def test_torch_morse_table_and_parameter_derivatives(tables, name):
    t = tables[name]
    s = _spec(t)
    idx, want = _golden(t, "v")
    f64 = dict(dtype=torch.float64)
    de = torch.tensor([s.de], requires_grad=True, **f64)
    a = torch.tensor([s.a], requires_grad=True, **f64)
    re = torch.tensor([s.re], requires_grad=True, **f64)
    r = uniform_grid(torch.tensor([s.r_min], **f64), torch.tensor([s.r_max], **f64), s.n)
    np.testing.assert_array_equal(r[0, idx].numpy(), s.grid()[idx])
    v = morse_table(r, de, a, re)[0, idx]
    _close(v.detach().numpy(), want, s.de)
    for k, key in enumerate(("dv_dde", "dv_da", "dv_dre")):
        _, dwant = _golden(t, key)
        got = np.array([torch.autograd.grad(v[j], (de, a, re), retain_graph=True)[k].item()
                        for j in range(len(idx))])
        np.testing.assert_allclose(got, dwant, rtol=1e-13, atol=64 * EPS * np.abs(dwant).max())


def test_batched_generator_is_stable():
    """The seeded batch (numpy default_rng(seed), draw order mu, D_e, a, r_e,
    N, M) reproduces these members and totals, so the CPU and GPU tests and
    the bench see the same bytes across library versions."""
    specs = batched(4096)
    members = {
        0: (3.7035034039127077, 23812.530083186535, 1.1187296906682986, 0.9128652703035955, 20363, 527),
        1: (1.1649998577106941, 33250.07525672095, 1.1427167929443516, 0.7539179827351236, 2805, 4044),
        2047: (46.986630660261724, 31018.871477521683, 1.704865289213889, 2.366896956872427, 6786, 271),
        4095: (10.04423140874866, 34466.09934760977, 1.221267523504044, 1.1230482696420196, 5811, 1327),
    }
    for i, want in members.items():
        s = specs[i]
        assert (s.mu, s.de, s.a, s.re, s.n, s.m) == want
        assert s.r_min == s.re - 0.6 and s.r_max == s.re + 6.0 / s.a
    assert sum(s.n for s in specs) == 72635671
    assert sum(s.m for s in specs) == 9286706
    assert BATCH_SEED == 20261202
