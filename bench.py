#!/usr/bin/env python3
"""Benchmark: fluctuating FCM mobility solves/s at 256^3 on one MI355X.

One step = one whole mobility solve of the 256^3 workload with 20,000 rods of
12 spheres (240,000 blobs, a = 1.86, dx = 1, mu = 1, kT = 1): the seeded
fluctuating-stress field and its interpolation, the right-hand side
[-V_rand; -F], and the right-preconditioned GMRES saddle-point solve to a
relative residual of 1e-8 -- timed for the FP32-bulk solve with FP64
iterative refinement (the headline `value`) and for the plain FP64 solve.
Inputs (geometry, forces, reference matrices) are resident in HBM before the
timed region; every step draws a fresh stress field from its own seed.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak, GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def percentile(xs, p):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * p / 100.0
    lo, hi = int(np.floor(k)), int(np.ceil(k))
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def pass_bytes(nx, ny, nz, nb, word, column=False):
    """Algorithmic HBM bytes per pass of one saddle mat-vec: S real words per
    field (S = nx*ny*2*(nz/2+1)); spread writes 3 fields, every FFT pass reads
    and writes 3 fields, interpolation reads 3 fields; per blob the spread
    reads position+force+permutation (52 B) and interpolation reads
    position+permutation and writes the velocity (52 B).  On the column path
    the z passes happen inside spreading (which then writes k-space rows) and
    interpolation (which reads them), so they move no bytes of their own."""
    S = nx * ny * 2 * (nz // 2 + 1)
    return {
        "spread": 3 * S * word + 52 * nb,
        "z_forward": 0 if column else 6 * S * word,
        "y_forward": 6 * S * word,
        "x_project": 6 * S * word,
        "y_inverse": 6 * S * word,
        "z_inverse": 0 if column else 6 * S * word,
        "interp": 3 * S * word + 52 * nb,
    }


def kernel_of_pass(t, n, dom):
    """Kernel behind profiled pass `dom` of the FP32 mobility product on an n^3
    grid under tuning `t` (as the library reports it: shapes resolved)."""
    p = 1  # FP32 slot
    cx, cp, yp, xp = t["col_cx"], t["col_points"][p], t["y_points"][p], t["x_points"][p]
    names = {"y_forward": f"cols_kernel<float, {n}, -1, false, {yp}>", "x_project": f"xproj_kernel<float, {n}, {xp}>",
             "y_inverse": f"cols_kernel<float, {n}, 1, false, {yp}>"}
    if t["column"][p] == 1:
        names.update({"spread": f"col_spread_kernel<float, {n}, {cx}, {cp}>",
                      "interp": f"col_interp_kernel<float, {n}, {cx}, {cp}>",
                      "interp_reduce": f"col_reduce_kernel<float, {cx}>"})
    else:
        names.update({"z_forward": f"zfwd_kernel<float, {n // 2}, false>", "z_inverse": f"zinv_kernel<float, {n // 2}>"})
    return names.get(dom, dom)


def measured_traffic(kernel, grid, nb):
    """HBM bytes per launch of `kernel` on this workload from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{grid}cubed_{nb}blobs.md")
    if not os.path.exists(path):
        return None
    section = None
    with open(path) as f:
        for line in f:
            if line.startswith("## "):
                section = line[3:].strip()
            elif line.startswith("traffic_bytes_per_launch:") and section and section in kernel:
                return int(line.split(":")[1])
    return None


def cpu_baseline(args, ref, x, q, F6, fp64_iters, Minv, N):
    """Times the CPU oracle (OpenMP over x slabs) on a bounded sample of the
    same workload (about 10 s of CPU work) and sums a solve from the measured
    parts: one fluctuating stress field (median of 3 seeds), one saddle-point
    mat-vec and one P^-1 at full size per iteration (medians of 9), and the
    modified Gram-Schmidt step of each iteration (gmres_solver.cpp:198-210),
    measured against 1, it/2 and it basis vectors and summed over the GPU FP64
    iteration count `it` (linear in j)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    n = args.grid
    oc = po.cfg(n, ng=7, dx=1.0, a=1.86, mu=1.0, kT=1.0, dt=args.dt, slip_channel=args.slip_channel)
    o = po.Fcm(oc, x.shape[0] * ref.shape[0])
    Y = po.fill_blob_pos(oc, ref, x, q)
    o.set_up_gaussians(Y)
    b = po.Bodies(ref, x, q, Minv, N)
    def timed(fn, reps):
        ts = []
        for r in range(reps):
            t0 = time.perf_counter()
            fn(r)
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts)

    def field(r):
        o.fill_stress(1 + r)
        o.random_field()
        o.random_interp()

    t_rand = timed(field, 3)
    v = np.random.default_rng(0).standard_normal(b.n)
    t_mv = timed(lambda r: po.saddle_matvec(o, b, v), 9)
    t_pinv = timed(lambda r: po.apply_Pinv(b, v), 9)
    o.close()
    it = max(2, fp64_iters)
    js = (1, it // 2, it)
    t_orth = [statistics.median(po.gmres_orth_seconds(b.n, j) for _ in range(3)) for j in js]
    slope = (t_orth[2] - t_orth[0]) / (it - 1)
    mgs = sum(t_orth[0] + slope * (j - 1) for j in range(1, it + 1))
    t_solve = t_rand + it * (t_mv + t_pinv) + mgs + t_pinv
    return {
        "value": 1.0 / t_solve,
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle (C restatement, FP64, {threads} OpenMP threads) at {n}^3 with {Y.size // 3} blobs, "
                   f"a solve summed from measured parts (medians): 1 fluctuating field {t_rand:.2f} s + {it} "
                   f"iterations x (saddle mat-vec {t_mv:.2f} s + P^-1 {t_pinv * 1e3:.1f} ms) + modified Gram-Schmidt "
                   f"(gmres_solver.cpp:198-210) measured at j = {js[0]}, {js[1]}, {js[2]} basis vectors "
                   f"({', '.join(f'{t * 1e3:.1f}' for t in t_orth)} ms) and summed over j = 1..{it}: "
                   f"{mgs:.2f} s; total {t_solve:.1f} s"),
    }


def fp64_true_residuals(s, seeds, args, F6, fa):
    """||b - A x||/||b|| of the FP64 solves of `seeds` with the FP64 operator
    (the plain solve stops on the Givens estimate): each seed's solve is run
    again after the timed region (bitwise the same solve) and one FP64
    saddle-point product is applied to its solution."""
    out = []
    for sd in seeds:
        rep = s.mobility_solve(sd, tol=args.tol, max_iter=args.max_iter, precision=fa.FCM_PREC_FP64)
        sol = s.solution()
        rhs = np.concatenate([-s.random_interp(), -np.asarray(F6).ravel()])
        out.append((rep["iterations"], float(np.linalg.norm(rhs - s.saddle_matvec(sol)) / np.linalg.norm(rhs))))
    return out


def bench_batch(args):
    """64 independent uneven instances (grids 32..128^3, 50-2000 rods of 6-12
    spheres, log-uniform) solved as one batch, against the same solves one at
    a time on the same GPU."""
    import math
    import timstochrigidbodiessim_amd as fa
    rng = np.random.default_rng(3)
    insts, tuned = [], []
    t0 = time.perf_counter()
    for i in range(args.instances):
        n = int(rng.choice([32, 48, 64, 96, 128]))
        NB = int(round(math.exp(rng.uniform(math.log(50), math.log(2000)))))
        nbpb = int(rng.integers(6, 13))
        cfg = fa.make_config(n, ng=7, dx=1.0, a=1.86, mu=1.0, kT=0.0, dt=args.dt, max_blobs=NB * nbpb,
                             max_bodies=NB, max_blobs_per_body=nbpb, max_krylov=args.max_iter, fp32=True,
                             random_field=False)
        ref, x, q, F6, nov = fa.synth_rods(100 + i, NB, nbpb, 1.86, [float(n)] * 3, max_attempts=5)
        s = fa.FcmSolver(cfg)
        s.set_geometry(ref)
        s.ref_matrices()
        s.set_bodies(x, q)
        if not args.no_autotune:  # one timing per shape class, the cached choice for the rest
            for prec in (fa.FCM_PREC_FP32, fa.FCM_PREC_FP64):  # FP32 solves refine with FP64 products
                t, hit = s.autotune_cached(prec, reps=3)
            if not hit:
                tuned.append({"grid": n, "blobs": NB * nbpb, "tuning": t})
        s.set_body_forces(F6)
        s.assemble_rhs(use_random=False)
        insts.append((s, n, NB, nbpb))
    log(f"[bench] batch set-up {time.perf_counter() - t0:.1f} s: {len(insts)} instances, "
        f"{sum(i[2] * i[3] for i in insts)} blobs in total, {len(tuned)} shape classes tuned")
    for tt in tuned:
        log(f"[bench] tuned {tt['grid']}^3, {tt['blobs']} blobs: {tt['tuning']}")
    prec = fa.FCM_PREC_FP32
    # one at a time
    t0 = time.perf_counter()
    alone = []
    for s, *_ in insts:
        _, r = s.solve(None, tol=args.tol, max_iter=args.max_iter, precision=prec, want_solution=False)
        alone.append(r)
    t_alone = time.perf_counter() - t0
    times = []
    for k in range(args.warmup + args.steps):
        with fa.FcmBatch(max_jobs=len(insts), worker_streams=args.workers, steps_per_round=args.steps_per_round) as b:
            t0 = time.perf_counter()
            for s, *_ in insts:
                b.submit(s, tol=args.tol, max_iter=args.max_iter, precision=prec)
            b.run()
            dt = time.perf_counter() - t0
            reps = [r for _, r in b.take()]
        if k >= args.warmup:
            times.append(dt)
        log(f"[bench] batch round {k}: {dt * 1e3:.1f} ms")
    t_batch = statistics.median(times)
    conv = sum(1 for r in reps if r["status"] == "converged")
    line = {
        "metric": "batched FCM mobility solves/s (64 uneven instances, FP32+refine, tol 1e-8)",
        "value": round(len(insts) / t_batch, 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t_batch * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 grid passes and inner Krylov basis + f64 coefficients/P^-1/refinement",
        "data": "synthetic (seeded rods and forces)",
        "config": {"workload": f"{len(insts)} instances, grids 32..128^3, 50-2000 rods x 6-12 spheres, kT=0",
                   "parallelism": "single GPU", "worker_streams": args.workers,
                   "steps_per_round": args.steps_per_round},
        "one_at_a_time_solves_per_s": round(len(insts) / t_alone, 3),
        "batch_speedup_vs_one_at_a_time": round(t_alone / t_batch, 3),
        "converged": conv, "iteration_cap": args.max_iter, "iterations_total": sum(r["iterations"] for r in reps),
        "tuned_shape_classes": tuned,
    }
    for s, *_ in insts:
        s.close()
    print(json.dumps(line), flush=True)


def bench_serve(args):
    """Poisson stream of mobility-solve requests (grids 64..256^3, 100-20,000
    rods of 10 spheres, kT = 1) served with continuous batching, FP32+refine
    and plain FP64."""
    from timstochrigidbodiessim_amd import serve
    import timstochrigidbodiessim_amd as fa
    out = {}
    for name, prec in (("fp32_refine", fa.FCM_PREC_FP32), ("fp64", fa.FCM_PREC_FP64)):
        reqs = serve.poisson_requests(rate=args.rate, count=args.requests, seed=11,
                                      grids=(64, 96, 128, 192, 256), rods=(100, 20000), nbpb=10,
                                      precision=prec, tol=args.tol, max_iter=args.max_iter)
        t0 = time.perf_counter()
        for r in reqs:
            r.make_geometry()
        log(f"[bench] {len(reqs)} requests generated ({time.perf_counter() - t0:.1f} s)")
        with serve.MobilityServer(max_jobs=args.max_jobs, worker_streams=args.workers,
                                  steps_per_round=args.steps_per_round, max_krylov=args.max_iter,
                                  autotune=not args.no_autotune) as srv:
            srv.prepare_refs(reqs)
            st = srv.run(reqs)
            st["handles_created"] = srv.created
            st["tunings_run"] = srv.tunings_run
            st["max_active"] = srv.max_active
        log(f"[bench] serve {name}: {st}")
        out[name] = st
    h = out["fp32_refine"]
    line = {
        "metric": "served FCM mobility solves/s (Poisson arrivals, 64..256^3, FP32+refine vs FP64); p50/p99 latency",
        "value": round(h["solves_per_s"], 3), "unit": "solves/s", "n_gpus": 1, "steps": args.requests,
        "warmup": 0, "ms_per_step": round(1e3 * h["wall_s"] / max(1, h["requests"]), 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 grid passes and inner Krylov basis + f64 coefficients/P^-1/refinement",
        "data": "synthetic (seeded Poisson stream, rods, forces and stress)",
        "config": {"workload": f"{args.requests} requests at {args.rate}/s, grids 64..256^3, 100-20000 rods x 10, "
                               f"kT=1, tol {args.tol:g}, iteration cap {args.max_iter}", "parallelism": "single GPU"},
        "converged": h["converged"],
        "p50_ms": round(1e3 * h["p50_latency_s"], 2), "p99_ms": round(1e3 * h["p99_latency_s"], 2),
        "fp32_refine": out["fp32_refine"], "fp64": out["fp64"],
    }
    print(json.dumps(line), flush=True)


def bench_timestep(args):
    """Config-5 time steps: 512^3 grid, 150,000 rods of 12 spheres (1.8M
    blobs), kT = 1, steric cutoff 2.2a, Euler-Maruyama steps each with a full
    fluctuating mobility solve (FP32+refine, tol 1e-8) and the body update.
    With --slip-channel: walls at y = 0 and y = grid/2, half the rods (the same
    volume fraction), wall springs, and a step that leaves a blob on or beyond a
    wall is undone and redone with a new seed (main.cpp:304-326)."""
    import timstochrigidbodiessim_amd as fa
    chan = args.slip_channel
    if args.ts_rods is None:
        args.ts_rods = 75000 if chan else 150000
    n, NB, nbpb, a = args.ts_grid, args.ts_rods, 12, 1.86
    L, margin = [float(n)] * 3, 0.0
    if chan:
        margin = a * (nbpb - 1) + 0.5  # every sphere between the walls
        L[1] = n / 2 - 2 * margin
    t0 = time.perf_counter()
    ref, x, q, _, nov = fa.synth_rods(5, NB, nbpb, a, L, max_attempts=args.place_attempts)
    x[:, 1] += margin
    nb = NB * nbpb
    phi = nb * 4.0 / 3.0 * np.pi * a ** 3 / (n ** 3 / (2 if chan else 1))
    log(f"[bench] timestep geometry {time.perf_counter() - t0:.1f} s: {NB} rods, phi {phi:.3f}, {nov} overlapping")
    cfg = fa.make_config(n, ng=7, dx=1.0, a=a, mu=1.0, kT=1.0, dt=args.dt, max_blobs=nb, max_bodies=NB,
                         max_blobs_per_body=nbpb, max_krylov=args.max_iter, fp32=True, random_field=True,
                         slip_channel=chan)
    log(f"[bench] footprint {fa.footprint(cfg) / 1e9:.1f} GB")
    s = fa.FcmSolver(cfg)
    s.set_geometry(ref)
    s.ref_matrices()
    s.set_bodies(x, q)
    s.set_body_forces(np.zeros((NB, 6)))
    cutoff = 2.2 * a
    times, reps, redone = [], [], 0
    for k in range(args.warmup + args.steps):
        t1 = time.perf_counter()
        if args.gdc:
            rep = s.gdc_step(1 + k, cutoff, fd_delta=1e-5, tol=args.tol, max_iter=args.max_iter,
                             precision=fa.FCM_PREC_FP32)
            redone += int(rep["redo"])
        else:
            rep = s.em_step(1 + k, cutoff, tol=args.tol, max_iter=args.max_iter, precision=fa.FCM_PREC_FP32)
        if chan and s.channel_wall_overlap():  # the step is undone and, as in main.cpp, taken again
            s.reset_to_start_of_step()
            redone += 1
        dt = time.perf_counter() - t1
        log(f"[bench] step {k}: {dt:.3f} s {rep}")
        if k >= args.warmup:
            times.append(dt)
            reps.append(rep)
    s.close()
    ms = 1e3 * statistics.mean(times)
    line = {
        "metric": "Brownian time steps/s at 512^3 with 1.8M spheres (FP32+refine GMRES at 1e-8, sterics, "
                  + ("drifter-corrector)" if args.gdc else "EM)") + (" (slip channel)" if chan else ""),
        "value": round(1e3 / ms, 4), "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 1), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 grid passes and inner Krylov basis + f64 coefficients/P^-1/refinement/constraints/quaternions",
        "data": "synthetic (seeded rods and stress stream)",
        "config": {"workload": f"{n}^3 grid, {NB} rods x {nbpb} spheres ({nb} blobs), a=1.86, kT=1, steric "
                               f"cutoff 2.2a, {'generalised drifter-corrector (divU by 6 finite differences)' if args.gdc else 'Euler-Maruyama'}, dt {args.dt:.5f}, volume fraction {phi:.3f}"
                               + (f", slip channel (walls at y = 0 and {n // 2}), {redone} step(s) undone for a "
                                  f"wall overlap" if chan else ""),
                   "parallelism": "single GPU"},
        "iterations": [r["iterations"] for r in reps], "status": sorted(set(r["status"] for r in reps)),
        "divU": [r["divU"] for r in reps] if args.gdc else None,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--workload", choices=["solve256", "batch", "serve", "timestep"], default="solve256",
                    help="solve256: the headline single-instance solve; batch: 64 uneven instances in one "
                         "batch; serve: a Poisson stream of requests with continuous batching; timestep: "
                         "Euler-Maruyama steps at 512^3 with 1.8M spheres")
    ap.add_argument("--ts-grid", type=int, default=512)
    ap.add_argument("--ts-rods", type=int, default=None, help="rods of the timestep workload: 150000, or 75000 "
                    "with --slip-channel")
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--steps-per-round", type=int, default=4)
    ap.add_argument("--requests", type=int, default=40)
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--max-jobs", type=int, default=16)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--rods", type=int, default=None,
                    help="rods of the solve256 workload: 20000, or 10000 with --slip-channel (the same volume "
                         "fraction in half the box)")
    ap.add_argument("--spheres", type=int, default=12)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--max-iter", type=int, default=None,
                    help="GMRES iteration cap (total over refinements) and Krylov capacity; default 600, "
                         "800 for batch and serve (instances near volume fraction 1 need 500-700 iterations)")
    ap.add_argument("--dt", type=float, default=np.pi * 1.86 ** 3 / 100.0)
    ap.add_argument("--place-attempts", type=int, default=20)
    ap.add_argument("--profile-reps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--gdc", action="store_true",
                    help="timestep: the generalised drifter-corrector step instead of Euler-Maruyama")
    ap.add_argument("--slip-channel", action="store_true",
                    help="solve256 / timestep in the slip-channel geometry: walls at y = 0 and y = grid/2, rods "
                         "inside")
    args = ap.parse_args()

    if args.max_iter is None:
        args.max_iter = 800 if args.workload in ("batch", "serve") else 600
    if args.rods is None:
        args.rods = 10000 if args.slip_channel else 20000
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the whole 256^3 grid lives on one MI355X and no collective or "
              "multi-GPU path exists in this build; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "timestep":
        return bench_timestep(args)

    import timstochrigidbodiessim_amd as fa

    n = args.grid
    L = [float(n)] * 3
    margin = 0.0
    if args.slip_channel:
        # rod centres at least one rod half-length (+0.5) from either wall, so
        # every sphere lies between y = 0 and y = n/2
        margin = 1.86 * (args.spheres - 1) + 0.5
        L[1] = n / 2 - 2 * margin
    t0 = time.perf_counter()
    ref, x, q, F6, nov = fa.synth_rods(2, args.rods, args.spheres, 1.86, L, max_attempts=args.place_attempts)
    x[:, 1] += margin
    nb = args.rods * args.spheres
    phi = nb * 4.0 / 3.0 * np.pi * 1.86 ** 3 / (n ** 3 / (2 if args.slip_channel else 1))
    log(f"[bench] synthetic rods: {args.rods} x {args.spheres}, volume fraction {phi:.3f}, "
        f"{nov} rods placed with overlap ({time.perf_counter() - t0:.1f} s)")
    cfg = fa.make_config(n, ng=7, dx=1.0, a=1.86, mu=1.0, kT=1.0, dt=args.dt, max_blobs=nb, max_bodies=args.rods,
                         max_blobs_per_body=args.spheres, max_krylov=args.max_iter, fp32=True, random_field=True,
                         slip_channel=args.slip_channel)
    s = fa.FcmSolver(cfg)
    s.set_geometry(ref)
    t0 = time.perf_counter()
    _, Minv, Nref = s.ref_matrices()
    log(f"[bench] reference matrices: {time.perf_counter() - t0:.2f} s")
    s.set_bodies(x, q)
    s.set_body_forces(F6)
    tuning = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        s.autotune(fa.FCM_PREC_FP32, reps=5)
        tuning = s.autotune(fa.FCM_PREC_FP64, reps=5)
        log(f"[bench] autotuned launch shapes {tuning} ({time.perf_counter() - t0:.2f} s)")

    results = {}
    modes = [("fp32_refine", fa.FCM_PREC_FP32)]
    if not args.skip_fp64:
        modes.append(("fp64", fa.FCM_PREC_FP64))
    for name, prec in modes:
        for w in range(args.warmup):
            rep = s.mobility_solve(1000 + w, tol=args.tol, max_iter=args.max_iter, precision=prec)
            log(f"[bench] warmup {name}: {rep}")
        s.L.fcm_sync(s.h)
        lat, reps = [], []
        t_start = time.perf_counter()
        for k in range(args.steps):
            t1 = time.perf_counter()
            rep = s.mobility_solve(k + 1, tol=args.tol, max_iter=args.max_iter, precision=prec)
            lat.append(time.perf_counter() - t1)
            reps.append(rep)
        s.L.fcm_sync(s.h)
        total = time.perf_counter() - t_start
        log(f"[bench] {name}: {args.steps} solves in {total:.3f} s; last {reps[-1]}")
        if prec == fa.FCM_PREC_FP64:
            tr = fp64_true_residuals(s, range(1, args.steps + 1), args, F6, fa)
            assert [i for i, _ in tr] == [r["iterations"] for r in reps], "FP64 re-run differs from the timed solve"
            for r, (_, res) in zip(reps, tr):
                r["true_rel_residual"] = res
            log(f"[bench] fp64 true residuals (untimed re-runs): max {max(t for _, t in tr):.3e}")
        results[name] = {
            "solves_per_s": args.steps / total,
            "ms_per_step": 1e3 * total / args.steps,
            "p50_ms": 1e3 * percentile(lat, 50),
            "p99_ms": 1e3 * percentile(lat, 99),
            "iterations": [r["iterations"] for r in reps],
            "refinements": [r["refinements"] for r in reps],
            "status": sorted(set(r["status"] for r in reps)),
            "true_rel_residual_max": max((r["true_rel_residual"] for r in reps), default=None),
        }

    # FP32+refine against the plain FP64 solve of the same stress seed
    agree = None
    if not args.skip_fp64:
        s.mobility_solve(1, tol=args.tol, max_iter=args.max_iter, precision=fa.FCM_PREC_FP64)
        sol64 = s.solution()
        s.mobility_solve(1, tol=args.tol, max_iter=args.max_iter, precision=fa.FCM_PREC_FP32)
        sol32 = s.solution()
        nl = 3 * nb
        agree = {
            "body_velocity_rel_l2": float(np.linalg.norm(sol32[nl:] - sol64[nl:]) / np.linalg.norm(sol64[nl:])),
            "blob_force_rel_l2": float(np.linalg.norm(sol32[:nl] - sol64[:nl]) / np.linalg.norm(sol64[:nl])),
        }
        log(f"[bench] fp32+refine vs fp64: {agree}")

    # roofline of the dominant kernel (per-pass device time, HIP events on the solver's stream)
    prof = {}
    column = s.tuning()["column"]
    for name, prec, word in (("fp32", fa.FCM_PREC_FP32, 4), ("fp64", fa.FCM_PREC_FP64, 8)):
        ms = s.profile_matvec(prec, args.profile_reps)
        by = pass_bytes(n, n, n, nb, word, column[prec] == 1)
        prof[name] = {k: {"ms": ms[k], "GBps": (by[k] / (ms[k] * 1e-3) / 1e9 if k in by and ms[k] > 0 else None)}
                      for k in ms if ms[k] > 0}
        log(f"[bench] passes {name}: " + ", ".join(f"{k} {v['ms']:.3f} ms" + (f" {v['GBps']:.0f} GB/s" if v['GBps'] else "")
                                                 for k, v in prof[name].items()))
    head = prof["fp32"]
    dom = max((k for k in head if k != "KT"), key=lambda k: head[k]["ms"])
    achieved = head[dom]["GBps"]
    kernel = kernel_of_pass(s.tuning(), n, dom)
    mv_bytes = sum(pass_bytes(n, n, n, nb, 4, column[fa.FCM_PREC_FP32] == 1).values())
    mv_ms = sum(v["ms"] for v in head.values())
    roofline = {"bound": "hbm", "kernel": kernel, "pass": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(kernel, n, args.rods * args.spheres),
                "algorithmic_bytes": pass_bytes(n, n, n, nb, 4, column[fa.FCM_PREC_FP32] == 1)[dom],
                "matvec_fp32": {"ms": round(mv_ms, 4), "GBps": round(mv_bytes / (mv_ms * 1e-3) / 1e9, 1)}}

    cpu = None
    if not args.no_cpu_baseline:
        try:
            it64 = int(np.median(results["fp64"]["iterations"])) if "fp64" in results else 50
            t0 = time.perf_counter()
            cpu = cpu_baseline(args, ref, x, q, F6, it64, Minv, Nref)
            log(f"[bench] cpu baseline ({time.perf_counter() - t0:.1f} s): {cpu}")
        except Exception as e:  # the baseline is reported, never required
            log(f"[bench] cpu baseline failed: {e!r}")

    head_r = results["fp32_refine"]
    line = {
        "metric": "FCM mobility solves/s at 256^3 (FP32+refine vs FP64) at 1e-8; p50/p99 latency"
                  + (" (slip channel)" if args.slip_channel else ""),
        "value": round(head_r["solves_per_s"], 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head_r["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 grid passes and inner Krylov basis + f64 coefficients/P^-1/refinement",
        "data": "synthetic (seeded rods, forces and stress stream)",
        "config": {"workload": f"{n}^3 grid, {args.rods} rods x {args.spheres} spheres ({nb} blobs), a=1.86, dx=1, "
                               f"mu=1, kT=1, ng=7, GMRES tol {args.tol:g}, volume fraction {phi:.3f}, "
                               f"{nov} rods overlapping"
                               + (f", slip channel (walls at y = 0 and y = {n // 2}, rod centres in "
                                  f"[{margin:.2f}, {n / 2 - margin:.2f}])" if args.slip_channel else ""),
                   "parallelism": "single GPU"},
        "p50_ms": round(head_r["p50_ms"], 3),
        "p99_ms": round(head_r["p99_ms"], 3),
        "fp32_refine": head_r,
        "fp64": results.get("fp64"),
        "speedup_fp32_refine_vs_fp64": (round(head_r["solves_per_s"] / results["fp64"]["solves_per_s"], 3)
                                        if "fp64" in results else None),
        "fp32_vs_fp64_solution": agree,
        "passes_ms": {k: {kk: round(vv["ms"], 4) for kk, vv in v.items()} for k, v in prof.items()},
        "roofline": roofline,
        "launch_tuning": tuning,
        "cpu_baseline": cpu,
    }
    s.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
