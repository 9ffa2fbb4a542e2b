#!/usr/bin/env python3
"""Benchmark of the Numerov-shooting hot path on one MI355X.

Workload (default ``batched``): 4,096 seeded Morse potentials (mu log-U[0.5,
100] amu, D_e U[5e3, 5e4] cm^-1, a U[1, 2.5] 1/A, r_e U[0.7, 3] A) with
ragged FP64 grids (log-U[2,000, 64,000] points on [r_e - 0.6, r_e + 6/a] A)
and trial counts (log-U[256, 8,192]), every bound level refined to 1e-10
relative, launch configuration from the autotuner.  A step is one pass of the
hot path over that input: Numerov scan of all trial energies, node-count
bracketing and refinement of every level, levels copied back.  The potential
tables are uploaded once before timing (inputs resident in HBM).  The run
also times the same solve done potential by potential (one context and one
launch sequence per potential) on a seeded sample of 64 potentials, and the
single heavy Morse potential below for a few steps.

``--workload single`` runs one heavy Morse potential (mu = 63.452 amu,
D_e = 12440 cm^-1, a = 1.8676 1/A, r_e = 2.666 A) on an FP64 grid of 200,000
points on r in [2, 8] A with 65,536 trial energies uniform in (min V, D_e).

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
dominant kernel (the scan, timed with HIP events on the stream it runs on),
the whole solve's FP64 fraction, and a CPU baseline (the C oracle on the host
cores, on a stated subset of the same workload, all cores and one core).
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


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_solve_work(oracle, prob, n_levels: int):
    """The oracle's scan + refinement of one potential as the GPU runs it on
    the batch (band windows, segmented shots of 64 segments): seconds,
    performed point-updates (counted as the library counts them) and
    converged levels."""
    t0 = time.perf_counter()
    nodes, far, resc, lv, it = oracle.solve(prob, 0, n_levels, 1e-10, 200, segmented=True, lanes=64, windows=True)
    dt = time.perf_counter() - t0
    win = prob.windows(True)
    steps = win[:, 1] - win[:, 0] - 1
    m = prob.m
    pu = 0.0
    for j0 in range(0, m, 64):
        pu += min(64, m - j0) * float(steps[oracle.group_band(j0, m, prob.e_lo, prob.e_top)])
    fe = oracle.band_far_exps(win)
    for v in range(n_levels):
        if it[v] > 0:
            br = oracle.bracket_level(nodes, far, resc, prob.e_lo, prob.e_top, v, fe, win)
            pu += float(it[v]) * float(steps[br.band])
    return dt, pu, int(np.sum(np.isfinite(lv)))


def cpu_baseline(all_cores_count: int = 256, one_core_count: int = 32) -> dict:
    """The C oracle (FP64, OpenMP over trial energies and levels) on subsets of
    the batched workload: its first ``all_cores_count`` potentials on every
    host core, its first ``one_core_count`` on one core."""
    import oracle
    from epseon_backend_amd.workloads import batched

    specs = batched(4096)
    threads = oracle.num_threads()

    def run(count: int, nthreads: int) -> tuple[float, float, int, int]:
        oracle.set_num_threads(nthreads)
        secs = pu = 0.0
        levels = 0
        for s in specs[:count]:
            prob = oracle.Problem(v=s.table(), r_min=s.r_min, r_max=s.r_max, mu=s.mu, e_top=s.de, m=s.m)
            if not prob.usable:
                continue
            dt, w, nl = _oracle_solve_work(oracle, prob, s.bound_levels() + 2)
            secs += dt
            pu += w
            levels += nl
        return secs, pu, levels, count

    s_all, pu_all, lv_all, n_all = run(all_cores_count, threads)
    s_one, pu_one, lv_one, n_one = run(one_core_count, 1)
    oracle.set_num_threads(threads)
    return {
        "value": pu_all / s_all,
        "unit": "point-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle scan + bracket + refinement (band windows, segmented shots) of the first {n_all} "
        f"of the 4096 batched potentials on {threads} threads ({pu_all:.3e} point-updates, {lv_all} levels "
        f"in {s_all:.2f} s), and of the first {n_one} on 1 thread ({pu_one:.3e} point-updates, {lv_one} "
        f"levels in {s_one:.2f} s)",
        "eigenlevels_per_s": lv_all / s_all,
        "single_core_value": pu_one / s_one,
        "single_core_eigenlevels_per_s": lv_one / s_one,
        "cpu_model": _cpu_model(),
    }


def per_potential_launches(capi, pots, rel_tol: float, sample: int = 256) -> dict:
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


def run_loads(capi, args) -> int:
    """Loads of the fitting batch with nothing else in flight (potential
    statistics, shifted tables, band windows and segment tables, host layout
    included): the load part of every forward pass of the fit, timed alone."""
    from epseon_backend_amd.workloads import DIFF_SEED, differentiation

    specs = differentiation(256)
    pots = [capi.Potential(values=s.table(), r_min=s.r_min, r_max=s.r_max, reduced_mass=s.mu, e_top=s.de,
                           n_energies=s.m, level_min=0, level_count=41) for s in specs]
    ctx = capi.Context(0)
    ms = []
    for i in range(args.loads_only + 1):
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.load(pots)
        ctx.synchronize()
        if i > 0:
            ms.append((time.perf_counter() - t0) * 1e3)
    ctx.close()
    out = {
        "workload": f"loads of the fitting batch: 256 Morse potentials (seed {DIFF_SEED}) x 50000 points, "
        "host V uploaded each time",
        "loads": args.loads_only,
        "load_ms": ms,
        "mean_load_ms": float(np.mean(ms)),
    }
    sys.stdout.write(json.dumps(out) + "\n")
    sys.stdout.flush()
    return 0


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


def load_traffic(workload: str) -> tuple[float | None, str | None]:
    """HBM bytes per scan launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration) in a separate profiled
    run, read from the committed summary: (bytes, file) or (None, None)."""
    path = REPO / "profiles" / f"scan_{workload}_traffic.json"
    if not path.exists():
        return None, None
    try:
        return float(json.loads(path.read_text())["bytes_per_launch"]), str(path.relative_to(REPO))
    except (KeyError, ValueError, OSError):
        return None, None


def single_summary(capi, rel_tol: float, warmup: int = 20, steps: int = 10) -> dict:
    """A few steps of the single heavy Morse workload (200,000 points, 65,536
    trial energies): its own context, timed like the main workload."""
    pots, config = single_workload(capi)
    ctx = capi.Context(0)
    ctx.load(pots)
    for _ in range(warmup):
        ctx.solve(rel_tol, 200)
        ctx.levels()
    pu = 0.0
    levels = 0
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.solve(rel_tol, 200)
        lv, _, _ = ctx.levels()
        st = ctx.stats()
        pu += st.scan_point_updates + st.refine_point_updates
        levels += int(st.levels_converged)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    from epseon_backend_amd.workloads import I2

    # Levels well inside the box against the analytic Morse levels (the
    # discretisation error at this step is about 1e-11 relative).
    ana = I2.analytic(np.arange(100))
    err = float(np.max(np.abs(lv[:100] - ana) / ana))
    ctx.scan()
    ctx.refine(rel_tol, 200)
    st = ctx.stats()
    ctx.close()
    scan_tflops = FLOP_PER_POINT_UPDATE * st.scan_point_updates / (st.scan_ms * 1e-3) / 1e12
    return {
        "workload": config["workload"],
        "ms_per_step": wall / steps * 1e3,
        "point_updates_per_s": pu / wall,
        "eigenlevels_per_s": levels / wall,
        "levels_per_step": levels // steps,
        "scan_ms": st.scan_ms,
        "refine_ms": st.refine_ms,
        "scan_frac": scan_tflops / FP64_VECTOR_PEAK_TFLOPS,
        "analytic_max_rel_err_v_lt_100": err,
        "check_ok": bool(err < 1e-9),
        "steps": steps,
        "warmup": warmup,
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps before timing (default: 20 single, 10 batched and fit, 0 stress; "
                         "the clock settles over the first ~40 ms of back-to-back solves)")
    ap.add_argument("--workload", choices=["single", "batched", "fit", "stress"], default="batched")
    ap.add_argument("--rel-tol", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="batched: keep the default launch configuration")
    ap.add_argument("--no-single", action="store_true", help="batched: skip the single heavy potential's steps")
    ap.add_argument("--loads-only", type=int, default=0, metavar="K",
                    help="fit: only load the fitting batch K times (after one untimed load) and report the "
                         "load times, each load's kernels running alone")
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
        return run_loads(capi, args) if args.loads_only > 0 else run_fit(capi, args)
    if args.workload == "stress":
        return run_stress(capi, args)
    pots, config = single_workload(capi) if args.workload == "single" else batched_workload(capi)
    ctx = capi.Context(0)
    ctx.load(pots)
    tuning = None
    if args.workload == "batched" and not args.no_autotune:
        tuning = ctx.autotune(args.rel_tol, 200).as_dict()

    # Result arrays reused by every step (levels(out=...)).
    n_levels = max(sum(p.level_count for p in pots), 1)
    out = (np.zeros(n_levels, np.float64), np.zeros(n_levels, np.int32), np.zeros(n_levels, np.int32))
    for _ in range(args.warmup):
        ctx.solve(args.rel_tol, 200)
        ctx.levels(out)

    solve_ms: list[float] = []
    solve_split = 0.0
    solve_graph = 0
    scan_pu = refine_pu = 0.0
    levels_ok = 0
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.solve(args.rel_tol, 200)
        lv, st, it = ctx.levels(out)
        s = ctx.stats()
        solve_ms.append(s.solve_ms)
        solve_split = s.solve_split
        solve_graph = int(s.solve_graph)
        scan_pu += s.scan_point_updates
        refine_pu += s.refine_point_updates
        levels_ok += int(s.levels_converged)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    # Sanity of the last step's levels: every level converged within the
    # iteration cap, and each potential's levels strictly increasing.
    check_ok = bool(np.all(st != capi.LEVEL_MAXITER))
    o = 0
    for p in pots:
        x = lv[o : o + p.level_count]
        x = x[np.isfinite(x)]
        check_ok = check_ok and bool(np.all(np.diff(x) > 0))
        o += p.level_count
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
    traffic, traffic_source = load_traffic(args.workload)
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
            "traffic": traffic,
            "kernel": "numerov_scan_kernel",
            "timing": "mean of 3 isolated scan launches, HIP events on the launch stream",
            "traffic_source": (f"{traffic_source}: rocprofv3 FETCH_SIZE + WRITE_SIZE of a separate profiled run "
                               "(committed constant, not measured in this run)") if traffic_source else None,
        },
    }
    mode = "segmented" if s.refine_mode == capi.REFINE_SEGMENTED else "lane"
    refine_achieved = (REFINE_FLOP_PER_POINT_UPDATE[mode] * (refine_pu / args.steps)
                       / (float(np.mean(refine_ms)) * 1e-3) / 1e12)
    out["refine_roofline"] = {
        "kernel": f"numerov_refine ({mode}, {int(s.refine_levels_per_block)} level(s) per block)",
        "bound": "valu",
        "achieved": refine_achieved,
        "peak": FP64_VECTOR_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": refine_achieved / FP64_VECTOR_PEAK_TFLOPS,
    }
    # The whole solve (scan, brackets, refinement; device time of the timed
    # solves): 7 FLOP per scan point-update, the refinement's per its mode.
    solve_flop = (FLOP_PER_POINT_UPDATE * scan_pu + REFINE_FLOP_PER_POINT_UPDATE[mode] * refine_pu) / args.steps
    solve_achieved = solve_flop / (float(np.mean(solve_ms)) * 1e-3) / 1e12
    out["solve_roofline"] = {
        "bound": "valu",
        "achieved": solve_achieved,
        "peak": FP64_VECTOR_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": solve_achieved / FP64_VECTOR_PEAK_TFLOPS,
    }
    out["scan_windows"] = int(s.scan_windows)
    out["levels_check_ok"] = check_ok
    if args.workload == "batched" and not args.no_single:
        out["single_heavy_morse"] = single_summary(capi, args.rel_tol)
    if not args.no_cpu_baseline and args.workload == "batched":
        out["cpu_baseline"] = cpu_baseline()
    if tuning is not None:
        out["tuning"] = tuning
    if args.workload == "batched":
        pp = per_potential_launches(capi, pots, args.rel_tol)
        out["per_potential_launches"] = pp
        out["batched_vs_per_potential"] = out["eigenlevels_per_s"] / pp["eigenlevels_per_s"]
    sys.stdout.write(json.dumps(out) + "\n")
    sys.stdout.flush()
    ok = check_ok and out.get("single_heavy_morse", {}).get("check_ok", True)
    if not ok:
        sys.stderr.write("bench.py: level checks failed (see levels_check_ok / single_heavy_morse.check_ok)\n")
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
