"""epseon_level_vjp_noalloc against epseon_level_vjp (device pointers).

The allocation-free entry point keeps the normalisation pass's segment maps
and shot checkpoints in the library's per-device scratch array.  Its
gradient must equal the ordinary entry point's bit for bit, on two calls in
a row (the second batch smaller, so the scratch holds the first call's
leftovers), and a batch needing more than the scratch maximum is rejected.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from epseon_backend_amd.workloads import I2, differentiation

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


def _context(capi, specs, levels):
    ctx = capi.Context(0)
    ctx.load([capi.Potential(values=s.table(), r_min=s.r_min, r_max=s.r_max, reduced_mass=s.mu, e_top=s.de,
                             n_energies=s.m, level_min=0, level_count=levels) for s in specs])
    ctx.solve(1e-10, 200)
    return ctx


def test_noalloc_vjp_equals_level_vjp_twice(capi):
    import torch

    cases = [(differentiation(12, n_points=6000), 20, 11), (differentiation(4, n_points=3000), 9, 12)]
    for k, (specs, levels, seed) in enumerate(cases):
        _log(f"case: call {k + 1}, {len(specs)} potentials x {specs[0].n} points, {levels} levels")
        with _context(capi, specs, levels) as ctx:
            n = sum(ctx.level_counts)
            g = torch.from_numpy(np.random.default_rng(seed).standard_normal(n)).cuda()
            # The allocation-free call first, on a context that has not
            # normalised yet, then the ordinary one.
            got = torch.full((ctx.v_span,), float("nan"), dtype=torch.float64, device="cuda")
            ctx.level_vjp_noalloc(g.data_ptr(), got.data_ptr())
            want = torch.full((ctx.v_span,), float("nan"), dtype=torch.float64, device="cuda")
            ctx.level_vjp_device(g.data_ptr(), want.data_ptr())
            a = got.cpu().numpy()
            b = want.cpu().numpy()
            assert np.all(np.isfinite(a))
            np.testing.assert_array_equal(a.view(np.int64), b.view(np.int64))
            assert np.abs(a).max() > 0.0
            # The context's own eigenfunction state was left alone: a second
            # ordinary call agrees as well.
            again = torch.zeros_like(want)
            ctx.level_vjp_device(g.data_ptr(), again.data_ptr())
            np.testing.assert_array_equal(again.cpu().numpy().view(np.int64), b.view(np.int64))


def test_noalloc_vjp_rejects_batches_above_the_scratch_maximum(capi):
    import torch

    # About 250 eigenfunction-shot segments per level on this grid: 3000
    # levels need about 84 MB of segment maps and checkpoints.
    _log("case: I2 200000-point grid with 3000 requested levels (over the scratch maximum)")
    with capi.Context(0) as ctx:
        ctx.load([capi.Potential(values=I2.table(), r_min=I2.r_min, r_max=I2.r_max, reduced_mass=I2.mu,
                                 e_top=I2.de, n_energies=4096, level_min=0, level_count=3000)])
        ctx.solve(1e-10, 200)
        g = torch.zeros(3000, dtype=torch.float64, device="cuda")
        out = torch.zeros(ctx.v_span, dtype=torch.float64, device="cuda")
        with pytest.raises(capi.EpseonError, match="maximum") as err:
            ctx.level_vjp_noalloc(g.data_ptr(), out.data_ptr())
        assert err.value.code == capi.EPSEON_ERR_INVALID
