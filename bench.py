#!/usr/bin/env python3
"""Benchmark of the Numerov-shooting hot path on one MI355X.

Workload (default ``single``): one heavy Morse potential (mu = 63.452 amu,
D_e = 12440 cm^-1, a = 1.8676 1/A, r_e = 2.666 A), an FP64 grid of 200,000
points on r in [2, 8] A, 65,536 trial energies uniform in (min V, D_e), every
bound level refined to 1e-10 relative.  A step is one pass of the hot path over
that input: Numerov scan of all trial energies, node-count bracketing and
refinement of every level, levels copied back.  The potential table is
uploaded once before timing (inputs resident in HBM).

``--workload batched`` runs the batched case instead: 4,096 seeded Morse
potentials with ragged grids (2k-64k points) and trial counts (256-8192); it
also times the same solve done potential by potential (one context and one
launch sequence per potential) on a seeded sample of 64 potentials.

``--workload stress`` runs 64 double-well potentials (Morse plus a Gaussian
barrier) on 1,000,000-point grids with 262,144 trial energies: a step is the
solve, every normalised eigenfunction written to device memory (one launch,
about 43 GB) and the adjoint pass with the eigenfunctions recomputed.

``--workload fit`` runs the differentiation case: 256 potentials drawn as in
the batched case (seed + 1) on 50,000-point grids, levels v = 0..40, a
least-squares loss against levels of +-2 % perturbed parameters plus
N(0, 0.01^2) noise; a step is forward (Morse tabulation in torch, solve)
plus backward (adjoint kernel, chain to D_e, a, r_e in torch), and the line
carries the grad/fwd wall-time ratio.

Prints one JSON line: metric/value (whole-job Numerov point-updates per
second over scan + refinement shots), eigenlevels/s, the roofline of the
dominant kernel (the scan, timed with HIP events on its own stream), and a
CPU baseline (the C oracle on the host cores, on a bounded sample of the same
workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

FP64_VECTOR_PEAK_TFLOPS = 78.6  # MI355X: 256 CU x 128 FP64 FLOP/clk x 2.4 GHz
# Algorithmic FLOP of one Numerov step (the count the workload's roofline is
# quoted in); the scan kernel issues 4 FP64 instructions for it (sub, mul,
# 2 FMA).
FLOP_PER_POINT_UPDATE = 7
# Refinement shots: a segmented shot carries two columns of the transfer
# matrix (10 FP64 instructions, 6 of them FMA, per grid point), a lane-mode
# shot one difference-form column (7, 3 of them FMA).
REFINE_FLOP_PER_POINT_UPDATE = {"segmented": 16, "lane": 10}
METRIC = "Numerov point-updates/s & eigenlevels/s at 1e-9 rel err; grad/fwd time ratio"

def single_workload(capi):
    from epseon_backend_amd.workloads import I2

    # Level count: every level below D_e (the analytic Morse count is 116).
    pots = [
        capi.Potential(
            values=I2.table(), r_min=I2.r_min, r_max=I2.r_max, reduced_mass=I2.mu, e_top=I2.de,
            n_energies=I2.m, level_min=0, level_count=120,
        )
    ]
    config = {
        "workload": "single heavy Morse potential (mu=63.452 amu, De=12440 cm-1, a=1.8676/A, "
        "re=2.666 A), FP64 grid of 200000 points on [2,8] A, 65536 trial energies in (min V, De), "
        "all bound levels refined to 1e-10 rel",
        "potentials": 1,
        "grid_points": I2.n,
        "trial_energies": I2.m,
    }
    return pots, config


def batched_workload(capi, count: int = 4096):
    from epseon_backend_amd.workloads import BATCH_SEED, batched

    specs = batched(count)
    pots = [
        capi.Potential(
            values=s.table(), r_min=s.r_min, r_max=s.r_max, reduced_mass=s.mu, e_top=s.de,
            n_energies=s.m, level_min=0, level_count=s.bound_levels() + 2,
        )
        for s in specs
    ]
    config = {
        "workload": f"batched: {count} seeded Morse potentials (seed {BATCH_SEED}), grids log-U[2000,64000] "
        "points on [re-0.6, re+6/a] A, trial energies log-U[256,8192], all bound levels to 1e-10 rel",
        "potentials": count,
        "grid_points": int(sum(s.n for s in specs)),
        "trial_energies": int(sum(s.m for s in specs)),
    }
    return pots, config


def cpu_baseline(target_seconds: float = 12.0) -> dict:
    """The C oracle (OpenMP over trial energies) on a bounded sample of the
    single-potential workload: the same grid with fewer trial energies."""
    import oracle
    from epseon_backend_amd.workloads import I2

    c = {"n": I2.n, "m": I2.m}
    prob = oracle.Problem(v=I2.table(), r_min=I2.r_min, r_max=I2.r_max, mu=I2.mu, e_top=I2.de, m=64)
    ut, b, stride = prob.pt, prob.beta, prob.stride
    threads = oracle.num_threads()
    m = 64 * threads
    t0 = time.perf_counter()
    oracle.scan(ut, b, prob.e_lo, prob.e_top, m, stride)
    dt = time.perf_counter() - t0
    m = int(min(c["m"], max(m, m * target_seconds / max(dt, 1e-3))))
    t0 = time.perf_counter()
    oracle.scan(ut, b, prob.e_lo, prob.e_top, m, stride)
    dt = time.perf_counter() - t0
    pu = (c["n"] - 2) * m
    return {
        "value": pu / dt,
        "unit": "point-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"Numerov scan of {m} of the 65536 trial energies over the full 200000-point grid "
        f"({pu:.3e} point-updates in {dt:.2f} s)",
    }


def per_potential_launches(capi, pots, rel_tol: float, sample: int = 64) -> dict:
    """The batched workload solved one potential at a time: each potential of
    a seeded sample in its own context (uploaded before timing), then a timed
    loop of solve + levels per potential."""
    rng = np.random.default_rng(1)
    idx = sorted(int(i) for i in rng.choice(len(pots), sample, replace=False))
    ctxs = []
    for i in idx:
        c = capi.Context(0, 8 << 20)
        c.load([pots[i]])
        c.solve(rel_tol, 200)
        c.levels()
        ctxs.append(c)
    levels = 0
    t0 = time.perf_counter()
    for c in ctxs:
        c.solve(rel_tol, 200)
        _, st, _ = c.levels()
        levels += int(np.sum(st == capi.LEVEL_OK))
    wall = time.perf_counter() - t0
    for c in ctxs:
        c.close()
    return {"eigenlevels_per_s": levels / wall, "potentials": sample, "levels": levels, "wall_s": wall}


def run_fit(capi, args) -> int:
    """Differentiation workload: forward + backward through the adjoint."""
    import torch

    from epseon_backend_amd.autograd import LevelSolver, LevelSpec, morse_table, numerov_levels, uniform_grid
    from epseon_backend_amd.workloads import DIFF_SEED, differentiation, perturbed

    levels = 41
    specs = differentiation(256)
    obs = perturbed(specs, seed=DIFF_SEED + 1)
    lspecs = [LevelSpec(s.r_min, s.r_max, s.mu, s.de, s.m, 0, levels) for s in specs]
    dev = torch.device("cuda:0")
    t = lambda x: torch.tensor(x, dtype=torch.float64, device=dev)  # noqa: E731
    r = uniform_grid(t([s.r_min for s in specs]), t([s.r_max for s in specs]), specs[0].n)
    solver = LevelSolver(0, rel_tol=args.rel_tol)
    with torch.no_grad():
        e_obs = numerov_levels(morse_table(r, t([s.de for s in obs]), t([s.a for s in obs]),
                                           t([s.re for s in obs])), lspecs, solver)
        gen = torch.Generator(device=dev).manual_seed(DIFF_SEED)
        e_obs = e_obs + 0.01 * torch.randn(e_obs.shape, dtype=torch.float64, device=dev, generator=gen)
    theta0 = [t([s.de for s in specs]), t([s.a for s in specs]), t([s.re for s in specs])]

    def step():
        th = [x.clone().requires_grad_() for x in theta0]
        torch.cuda.synchronize()
        a0 = time.perf_counter()
        e = numerov_levels(morse_table(r, *th), lspecs, solver)
        ok = torch.isfinite(e) & torch.isfinite(e_obs)
        loss = torch.where(ok, (e - e_obs) ** 2, torch.zeros_like(e)).sum()
        torch.cuda.synchronize()
        a1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        a2 = time.perf_counter()
        return a1 - a0, a2 - a1, int(ok.sum())

    for _ in range(args.warmup):
        step()
    fwd, bwd = [], []
    pu = 0.0
    levels_ok = 0
    solve_ms, vjp_ms, norm_ms = [], [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        f, b, n_ok = step()
        fwd.append(f)
        bwd.append(b)
        st = solver.last_stats
        pu += st.scan_point_updates + st.refine_point_updates + st.normalize_point_updates + st.vjp_point_updates
        solve_ms.append(st.solve_ms)
        vjp_ms.append(st.vjp_ms)
        norm_ms.append(st.normalize_ms)
        levels_ok += n_ok
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    solver.close()
    n_points = int(sum(s.n for s in specs))
    out = {
        "metric": METRIC,
        "value": pu / wall,
        "unit": "point-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"fit: 256 Morse potentials drawn as the batched case (seed {DIFF_SEED}) on "
            "50000-point grids, levels v=0..40, loss sum (E - E_obs)^2 against +-2% perturbed parameters "
            "+ N(0, 0.01^2); forward (tabulate, solve) + backward (adjoint kernel, chain to De, a, re)",
            "potentials": len(specs),
            "grid_points": n_points,
        },
        "grad_fwd_ratio": float(np.mean(bwd) / np.mean(fwd)),
        "forward_ms": float(np.mean(fwd)) * 1e3,
        "backward_ms": float(np.mean(bwd)) * 1e3,
        "solve_ms": float(np.mean(solve_ms)),
        "normalize_ms": float(np.mean(norm_ms)),
        "vjp_ms": float(np.mean(vjp_ms)),
        "eigenlevels_per_s": levels_ok / wall,
        "levels_per_step": levels_ok // args.steps,
    }
    sys.stdout.write(json.dumps(out) + "\n")
    sys.stdout.flush()
    return 0


def run_stress(capi, args) -> int:
    """Double-well stress workload: solve + eigenfunction output + adjoint."""
    import torch

    from epseon_backend_amd.workloads import STRESS_SEED, double_wells

    specs = double_wells(64)
    counts = [s.bound_levels() + 2 for s in specs]
    pots = [
        capi.Potential(values=s.table(), r_min=s.r_min, r_max=s.r_max, reduced_mass=s.mu, e_top=s.de,
                       n_energies=s.m, level_min=0, level_count=c)
        for s, c in zip(specs, counts)
    ]
    ctx = capi.Context(0)
    ctx.load(pots)
    dev = torch.device("cuda:0")
    psi_buf = torch.empty(ctx.eigenfunction_size(), dtype=torch.float64, device=dev)
    grad_v = torch.empty(ctx.v_span, dtype=torch.float64, device=dev)
    g = torch.ones(sum(counts), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def step():
        a0 = time.perf_counter()
        ctx.solve(args.rel_tol, 200)
        _, st, _ = ctx.levels()
        a1 = time.perf_counter()
        ctx.all_wavefunctions_device(psi_buf.data_ptr())
        a2 = time.perf_counter()
        ctx.level_vjp_device(g.data_ptr(), grad_v.data_ptr())
        a3 = time.perf_counter()
        return a1 - a0, a2 - a1, a3 - a2, int(np.sum(st == capi.LEVEL_OK))

    for _ in range(args.warmup):
        step()
    solve_t, psi_t, vjp_t = [], [], []
    pu = 0.0
    levels_ok = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a, b, c, n_ok = step()
        solve_t.append(a)
        psi_t.append(b)
        vjp_t.append(c)
        st = ctx.stats()
        # Eigenfunction output shoots every converged level once more (same
        # count as the normalisation pass).
        pu += (st.scan_point_updates + st.refine_point_updates + 2 * st.normalize_point_updates
               + st.vjp_point_updates)
        levels_ok += n_ok
    wall = time.perf_counter() - t0
    ctx.close()
    out = {
        "metric": METRIC,
        "value": pu / wall,
        "unit": "point-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"stress: 64 double-well potentials (Morse + B exp(-((r-r_b)/w)^2), seed {STRESS_SEED}), "
            "1000000-point grids, 262144 trial energies, all bound levels to 1e-10 rel, every eigenfunction "
            "written to device memory, adjoint pass with eigenfunctions recomputed",
            "potentials": len(specs),
            "grid_points": int(sum(s.n for s in specs)),
            "trial_energies": int(sum(s.m for s in specs)),
        },
        "solve_s": float(np.mean(solve_t)),
        "eigenfunctions_s": float(np.mean(psi_t)),
        "adjoint_s": float(np.mean(vjp_t)),
        "grad_fwd_ratio": float(np.mean(vjp_t) / np.mean(np.add(solve_t, psi_t))),
        "eigenfunction_bytes_per_step": int(sum(c * s.n for s, c in zip(specs, counts)) * 8),
        "eigenlevels_per_s": levels_ok / wall,
        "levels_per_step": levels_ok // args.steps,
    }
    sys.stdout.write(json.dumps(out) + "\n")
    sys.stdout.flush()
    return 0


def load_traffic(workload: str) -> float | None:
    """HBM bytes per scan launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), if committed."""
    path = REPO / "profiles" / f"scan_{workload}_traffic.json"
    if not path.exists():
        return None
    try:
        return float(json.loads(path.read_text())["bytes_per_launch"])
    except (KeyError, ValueError, OSError):
        return None


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps before timing (default: 20 single, 10 batched and fit, 0 stress; "
                         "the clock settles over the first ~40 ms of back-to-back solves)")
    ap.add_argument("--workload", choices=["single", "batched", "fit", "stress"], default="single")
    ap.add_argument("--rel-tol", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="batched: keep the default launch configuration")
    ap.add_argument(
        "--copy-calibration",
        action="store_true",
        help="only run a plain 1 GiB device-to-device copy 5 times (the known-size kernel whose "
        "FETCH_SIZE/WRITE_SIZE counters calibrate the traffic of the scan)",
    )
    args = ap.parse_args()
    if args.warmup is None:
        args.warmup = {"single": 20, "batched": 10, "fit": 10, "stress": 0}[args.workload]

    if args.copy_calibration:
        import torch

        n = (1 << 30) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        sys.stdout.write(json.dumps({"copy_calibration_bytes_read": n * 8, "bytes_written": n * 8, "launches": 5}) + "\n")
        return 0

    if args.gpus != 1:
        sys.stderr.write(
            "bench.py measures one GPU only: the Numerov path shards potentials (or trial "
            "energies) across GPUs with no exchange, and this build runs on one MI355X. "
            f"--gpus {args.gpus} is not supported.\n"
        )
        return 2

    from epseon_backend_amd import capi

    if capi.device_count() < 1:
        sys.stderr.write("no HIP device visible\n")
        return 1

    if args.workload == "fit":
        return run_fit(capi, args)
    if args.workload == "stress":
        return run_stress(capi, args)
    pots, config = single_workload(capi) if args.workload == "single" else batched_workload(capi)
    ctx = capi.Context(0)
    ctx.load(pots)
    tuning = None
    if args.workload == "batched" and not args.no_autotune:
        tuning = ctx.autotune(args.rel_tol, 200).as_dict()

    for _ in range(args.warmup):
        ctx.solve(args.rel_tol, 200)
        ctx.levels()

    solve_ms: list[float] = []
    solve_split = 0.0
    solve_graph = 0
    scan_pu = refine_pu = 0.0
    levels_ok = 0
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.solve(args.rel_tol, 200)
        lv, st, it = ctx.levels()
        s = ctx.stats()
        solve_ms.append(s.solve_ms)
        solve_split = s.solve_split
        solve_graph = int(s.solve_graph)
        scan_pu += s.scan_point_updates
        refine_pu += s.refine_point_updates
        levels_ok += int(s.levels_converged)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    # Roofline of the scan kernel: isolated scans (one launch each), timed
    # with HIP events on the context's stream; the solve above may overlap
    # scan and refinement on two streams.
    iso_scan_ms, iso_refine_ms = [], []
    for _ in range(3):
        ctx.scan()
        ctx.refine(args.rel_tol, 200)
        s = ctx.stats()
        iso_scan_ms.append(s.scan_ms)
        iso_refine_ms.append(s.refine_ms)
    ctx.close()

    ms_per_step = wall / args.steps * 1e3
    value = (scan_pu + refine_pu) / wall
    scan_ms = iso_scan_ms
    refine_ms = iso_refine_ms
    avg_scan_s = float(np.mean(scan_ms)) * 1e-3
    achieved = FLOP_PER_POINT_UPDATE * (scan_pu / args.steps) / avg_scan_s / 1e12
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "point-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": config,
        "eigenlevels_per_s": levels_ok / wall,
        "levels_per_step": levels_ok // args.steps,
        "solve_ms": float(np.mean(solve_ms)),
        "solve_split": solve_split,
        "solve_graph": solve_graph,
        "scan_ms": float(np.mean(scan_ms)),
        "refine_ms": float(np.mean(refine_ms)),
        "scan_point_updates_per_step": scan_pu / args.steps,
        "refine_point_updates_per_step": refine_pu / args.steps,
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP64_VECTOR_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_VECTOR_PEAK_TFLOPS,
            "traffic": load_traffic(args.workload),
        },
    }
    mode = "segmented" if s.refine_mode == capi.REFINE_SEGMENTED else "lane"
    refine_achieved = (REFINE_FLOP_PER_POINT_UPDATE[mode] * (refine_pu / args.steps)
                       / (float(np.mean(refine_ms)) * 1e-3) / 1e12)
    out["refine_roofline"] = {
        "kernel": f"numerov_refine ({mode})",
        "bound": "valu",
        "achieved": refine_achieved,
        "peak": FP64_VECTOR_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": refine_achieved / FP64_VECTOR_PEAK_TFLOPS,
    }
    out["scan_windows"] = int(s.scan_windows)
    if not args.no_cpu_baseline and args.workload == "single":
        out["cpu_baseline"] = cpu_baseline()
    if tuning is not None:
        out["tuning"] = tuning
    if args.workload == "batched":
        pp = per_potential_launches(capi, pots, args.rel_tol)
        out["per_potential_launches"] = pp
        out["batched_vs_per_potential"] = out["eigenlevels_per_s"] / pp["eigenlevels_per_s"]
    sys.stdout.write(json.dumps(out) + "\n")
    sys.stdout.flush()
    return 0


if __name__ == "__main__":
    sys.exit(main())
