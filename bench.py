"""Benchmark of the CCD++ hot path on one MI355X.

Workload: the 480,189 x 17,770 matrix with 99,072,112 training and
1,408,395 test ratings (seeded synthetic, power-law degrees), k=40,
lambda=0.05, 5 inner iterations.  One step = one outer iteration: all 40
rank-one updates with 5 inner sweeps each over the whole matrix, plus the
residual updates and the device test RMSE.  The default --steps 10 is the
configuration's T=10 solve, timed from the reference initial state
(W from initial_col, H = 0) with the data already resident in HBM.

Prints one JSON line (metric, value = nnz*k updates/s, roofline of the
dominant kernel measured with HIP events inside the timed region,
cpu_baseline = the OpenMP oracle on a bounded sample of the same workload).

usage: python bench.py [--gpus 1] [--steps 10] [--warmup 1] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD public spec
DEFAULT_VARIANT = 3  # tiled sweeps, sigma 16: the fastest variant on this workload


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ds, W0, k, lam, inner, budget_s=15.0):
    """OpenMP oracle (the restated cdmf_ref loops, schedule(dynamic,500)) on
    the same matrix, one OpenMP thread per host core: whole ranks of
    steady-state outer iterations (add-back, 5 inner sweeps, subtract; outer
    iterations 2, 3, ...), as many as fit in `budget_s`.  R's values and W,
    H are copied once, before the clock starts; the ranks then carry the
    residual from one to the next in place, as the solver does."""
    import numpy as np

    import oracle
    threads = os.cpu_count() or 1
    H0 = np.zeros((k, ds.n), np.float32)
    st = oracle.RankState(ds, W0, H0)
    ranks, oiter = 0, 2
    t0 = time.perf_counter()
    while True:
        oracle.ccd_ranks(st, ranks % k, ranks % k + 1, lam, oiter, inner, threads=threads)
        ranks += 1
        if ranks % k == 0:
            oiter += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {
        "value": ranks * ds.nnz / dt,
        "unit": "nnz*k updates/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": cpu_model(),
        "sample": "%d ranks (outer iterations 2.., %d ranks each: add-back, %d inner sweeps, subtract) on the "
                  "full %dx%d matrix, residual carried from rank to rank, %.1f s, %d OpenMP threads "
                  "(os.cpu_count()) on %s"
                  % (ranks, k, inner, ds.m, ds.n, dt, threads, cpu_model()),
    }


def reference_format_bytes(kind, m, n, nnz):
    """Algorithmic HBM bytes of one launch in the reference's own format (u32
    indices, f32 values, f32 vectors): the byte model of the path.
    sweep over the CSC copy (v update): 8 nnz + 4(n+1) + 4m + 4n;
    over the CSR copy (u update): 8 nnz + 4(m+1) + 4n + 4m.  A fused pass
    (residual update and first sweep of a rank in one pass over the H-first
    layouts) adds the residual write (4 nnz) and one read each of the
    previous rank's gathered and per-row vectors."""
    if kind == "sweep_f":
        return 8 * nnz + 4 * (n + 1) + 4 * m + 4 * n
    if kind == "sweep_s":
        return 8 * nnz + 4 * (m + 1) + 4 * n + 4 * m
    if kind == "fused_f":
        return 12 * nnz + 4 * (n + 1) + 2 * 4 * m + 2 * 4 * n + 4 * n
    if kind == "fused_s":
        return 12 * nnz + 4 * (m + 1) + 3 * 4 * n + 2 * 4 * m + 4 * m
    raise ValueError(kind)


def load_traffic():
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--small", action="store_true", help="the 2,000 x 1,000 case instead (quick check)")
    ap.add_argument("--variant", type=int, default=DEFAULT_VARIANT, help="1/2 = degree-binned, 3 = tiled sweeps")
    ap.add_argument("--sigma", type=int, default=16, help="elements per lane of a tile (tiled sweeps)")
    ap.add_argument("--split-rows", type=int, default=1, choices=(0, 1),
                    help="tiled sweeps: keep the elements gathered from beyond the LDS prefix in tiles of their own")
    ap.add_argument("--graph", type=int, default=1, choices=(0, 1),
                    help="replay outer iterations 2.. from one captured HIP graph (bit-identical to eager "
                         "launches; the dominant kernel is timed on outer iteration 1, which runs eagerly)")
    ap.add_argument("--blocks-per-cu", type=int, default=0,
                    help="persistent sweep blocks per CU (0: as many as the kernel's registers allow)")
    ap.add_argument("--batch", action="store_true",
                    help="the batched workload instead: 64 independent problems in one persistent launch, "
                         "against 64 sequential single-problem runs")
    ap.add_argument("--batch-scale", type=float, default=1.0, help="shrink the batched problems (quick checks)")
    ap.add_argument("--foldin", action="store_true",
                    help="the held-out-users fold-in study on the large shape (plain vs learned warm start)")
    ap.add_argument("--foldin-users", type=int, default=0, help="rows of the fold-in study's matrix (0: 480,189)")
    ap.add_argument("--foldin-tol", type=float, default=1e-4,
                    help="fold-in study: distance to the converged test RMSE that counts as reached")
    ap.add_argument("--foldin-max-sweeps", type=int, default=60, help="fold-in study: longest sweep curve recorded")
    ap.add_argument("--probe", action="store_true",
                    help="run a 256 MiB stream-copy kernel first (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the solver has no multi-GPU path; use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from cdmf_amd import data as D
    from cdmf_amd import solver as S

    if S.device_count() < 1:
        sys.exit("bench.py: no HIP device visible")
    if args.batch:
        bench_batch(args)
        return
    if args.foldin:
        bench_foldin(args)
        return

    if args.small:
        cfg, spec = D.SMALL_SOLVE, D.SMALL
        t = time.perf_counter()
        ds = D.small()
    else:
        cfg, spec = D.NETFLIX_SOLVE, D.NETFLIX
        t = time.perf_counter()
        ds = D.netflix()
    log("generated %dx%d nnz=%d test=%d in %.1fs" % (ds.m, ds.n, ds.nnz, ds.nnz_test, time.perf_counter() - t))
    k, lam, inner = cfg.k, cfg.lam, cfg.inner
    W0 = D.initial_col(k, ds.m)
    H0 = np.zeros((k, ds.n), np.float32)

    s = S.Solver(0)
    s.set_plan(sigma=args.sigma, blocks_per_cu=args.blocks_per_cu, split_rows=args.split_rows)
    if args.probe:
        ms = s.copy_probe(256 << 20, reps=5)
        log("copy probe: 256 MiB in %.1f us (%.0f GB/s read+write)" % (ms * 1e3, 2 * (256 << 20) / (ms * 1e-3) / 1e9))
    t = time.perf_counter()
    s.upload(ds)
    upload_s = time.perf_counter() - t
    log("uploaded in %.1fs" % upload_s)

    # The warmup times every kernel class (event records between kernels add
    # a few microseconds each); the timed run times only the dominant sweep
    # kernel, with events bound to its own dispatches, which add nothing
    # between kernels.  A bound pair still costs about 5 us of device time,
    # so the library times one launch in 5 of that kernel (the launches are
    # identical passes over the same layout).
    sweep_kinds = ("sweep_f", "sweep_s", "fused_f", "fused_s")
    all_stats = None
    timed_kinds = sweep_kinds
    if args.warmup > 0:
        s.set_factors(W0, H0)
        t = time.perf_counter()
        s.run(k, lam, args.warmup, inner, variant=args.variant, profile=True)
        all_stats = s.kernel_stats()
        timed_kinds = (max(sweep_kinds, key=lambda kk: all_stats[kk]["total_ms"]),)
        if args.graph:  # capture the replayed outer iteration before the clock starts
            s.set_factors(W0, H0)
            s.run(k, lam, 2, inner, variant=args.variant, use_graph=True)
        log("warmup %d outer iteration(s) in %.2fs" % (args.warmup, time.perf_counter() - t))

    s.set_factors(W0, H0)
    t0 = time.perf_counter()
    res = s.run(k, lam, args.steps, inner, variant=args.variant, profile=True, use_graph=bool(args.graph),
                profile_kinds=timed_kinds)  # returns after the stream drains
    t1 = time.perf_counter()
    wall = t1 - t0
    log("timed %d outer iterations in %.3fs; test RMSE %s" % (args.steps, wall, ["%.6f" % r for r in res.rmse]))

    stats = s.kernel_stats()
    dom = max(timed_kinds, key=lambda kk: stats[kk]["total_ms"])
    d = stats[dom]
    if all_stats is None:
        all_stats = stats
    avg_ms = d["total_ms"] / max(d["launches"], 1)
    # H-first: the first-swept layout is the CSC copy
    model_bytes = reference_format_bytes(dom, ds.m, ds.n, ds.nnz)
    achieved = model_bytes / (avg_ms * 1e-3) / 1e9
    achieved_fmt = d["bytes_per_launch"] / (avg_ms * 1e-3) / 1e9
    kernels = {kk: dict(avg_us=1e3 * v["total_ms"] / max(v["launches"], 1), launches=v["launches"],
                        gbs=(v["bytes_per_launch"] / (v["total_ms"] / max(v["launches"], 1) * 1e-3) / 1e9)
                        if v["launches"] else 0.0)
               for kk, v in all_stats.items() if v["launches"]}
    traffic = load_traffic()
    roofline = {
        "bound": "hbm",
        "kernel": ("tiled " if args.variant == 3 else "binned ") + {
            "sweep_f": "sweep (CSC, v update)", "sweep_s": "sweep (CSR, u update)",
            "fused_f": "fused residual+sweep (CSC)", "fused_s": "fused residual+sweep (CSR)"}[dom],
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic.get("v%d" % args.variant, {}).get(dom, {}).get("hbm_bytes_per_launch") if traffic else None,
        "algorithmic_bytes_per_launch": model_bytes,
        "algorithmic_bytes_model": "u32 indices + f32 values + f32 vectors (the reference's format)",
        "format_bytes_per_launch": d["bytes_per_launch"],
        "achieved_format_bytes": round(achieved_fmt, 1),
        "frac_format_bytes": round(achieved_fmt / HBM_PEAK_GBS, 4),
        "avg_launch_us": round(avg_ms * 1e3, 2),
    }
    value = args.steps * k * ds.nnz / wall
    out = {
        "metric": "CCD++ nnz*k updates/s",
        "value": value,
        "unit": "nnz*k updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic, seeded power-law ratings (%d x %d, %d train / %d test)" % (ds.m, ds.n, ds.nnz, ds.nnz_test),
        "config": {"workload": "CCD++ rank-one loop, %d x %d, nnz=%d, k=%d, lambda=%g, %d outer x %d inner"
                               % (ds.m, ds.n, ds.nnz, k, lam, args.steps, inner),
                   "m": ds.m, "n": ds.n, "nnz": ds.nnz, "nnz_test": ds.nnz_test, "k": k, "lambda": lam,
                   "outer": args.steps, "inner": inner, "seed": spec["seed"], "variant": args.variant,
                   "row_degree_exponent": round(D.row_degree_exponent(**spec), 4) if "max_row_degree" in spec
                   else None,
                   "row_degree_rule": "power law on [1, 17000], exponent fitted to the mean degree" if
                   "max_row_degree" in spec else "uniform positions",
                   "sigma": args.sigma if args.variant == 3 else None, "graph": bool(args.graph)},
        "roofline": roofline,
        "test_rmse": res.rmse,
        "loop_time_s": s.loop_time(),
        # ccd_upload from host buffers (degree renumbering, row sorting and
        # both compressed copies over PCIe): outside `value`, which starts
        # with the ratings resident in HBM
        "upload_s": upload_s,
        "kernels_warmup": kernels,
    }
    s.close()
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(ds, W0, k, lam, inner)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def bench_batch(args):
    """64 independent factorizations (shapes, k and lambda drawn per problem,
    8 outer x 5 inner, FP32): the sequential baseline runs them one after the
    other with the single-problem tiled path; the batched run solves all of
    them in one persistent launch with the autotuned plan, eagerly and as a
    graph replay.  Every timed run starts from the same initial factors."""
    import numpy as np

    from cdmf_amd import data as D
    from cdmf_amd import solver as S

    specs = D.batch_specs(scale=args.batch_scale)
    t = time.perf_counter()
    solvers, inits = [], []
    for bp in specs:
        ds = D.batch_dataset(bp)
        s = S.Solver(0)
        s.set_plan(sigma=args.sigma, split_rows=0)  # the batched kernel's plain tile layouts
        s.upload(ds)
        W0 = D.initial_col(bp.solve.k, ds.m)
        H0 = np.zeros((bp.solve.k, ds.n), np.float32)
        solvers.append(s)
        inits.append((W0, H0, ds.nnz))
    log("generated and uploaded %d problems in %.1fs" % (len(specs), time.perf_counter() - t))
    updates = sum(bp.solve.outer * bp.solve.k * nnz for bp, (_, _, nnz) in zip(specs, inits))

    def reset():
        for s, (W0, H0, _) in zip(solvers, inits):
            s.set_factors(W0, H0)

    # sequential: one ccd_run per problem
    per_problem = []

    def sequential():
        tot = 0.0
        per_problem.clear()
        for s, bp in zip(solvers, specs):
            t0 = time.perf_counter()
            s.run(bp.solve.k, bp.solve.lam, bp.solve.outer, bp.solve.inner, variant=S.VARIANT_TILED,
                  rmse_every=False)
            dt = time.perf_counter() - t0
            per_problem.append(dt)
            tot += dt
        return tot

    reset()
    sequential()  # warm-up (tiling, first launches)
    reset()
    seq_s = sequential()
    seq_rmse = [s.rmse() for s in solvers]
    log("sequential: %.3f s" % seq_s)

    b = S.Batch(0)
    reset()
    for s, bp in zip(solvers, specs):
        b.add(s, bp.solve.k, bp.solve.lam, bp.solve.outer, bp.solve.inner)
    reset()
    trial_ms = b.tune()
    plan = b.plan()
    tune_times = b.tune_times()
    n_solo = min(3, len(specs) - 1) + 1
    tune_solo, tune_deal = tune_times[:n_solo], tune_times[n_solo:]
    log("autotuned plan %s (%.2f ms); whole runs with 0, 1, ... problems solved alone: %s ms; with the tile deal "
        "at 8, 32, never (16 above): %s ms" % (plan, trial_ms, ["%.1f" % x for x in tune_solo],
                                               ["%.1f" % x for x in tune_deal]))
    reset()
    b.run()  # warm-up
    reset()
    t0 = time.perf_counter()
    b.run()
    eager_wall = time.perf_counter() - t0
    reset()
    b.run(use_graph=True)  # capture
    reset()
    t0 = time.perf_counter()
    graph_dev = b.run(use_graph=True)
    graph_wall = time.perf_counter() - t0
    rmse = [s.rmse() for s in solvers]
    teams = b.teams()
    b.close()
    for s in solvers:
        s.close()
    worst = max(abs(a - c) for a, c in zip(rmse, seq_rmse))
    out = {
        "metric": "batched CCD++ nnz*k updates/s (aggregate)",
        "value": updates / graph_wall,
        "unit": "nnz*k updates/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "higher_is_better": True,
        "dtype": "fp32",
        "data": "synthetic, %d seeded power-law problems (%s total ratings)" % (len(specs), sum(x[2] for x in inits)),
        "config": {"workload": "batched CCD++: %d problems, rows log-U[1k,200k], cols log-U[500,20k], nnz log-U[10k,5M], "
                               "k in {8,16,32,64}, lambda in {.01,.05,.1}, 8 outer x 5 inner" % len(specs),
                   "scale": args.batch_scale, "plan": plan},
        "batched_graph_s": graph_wall,
        "batched_graph_device_s": graph_dev,
        "batched_eager_s": eager_wall,
        "sequential_s": seq_s,
        "sequential_updates_per_s": updates / seq_s,
        "speedup_vs_sequential": seq_s / graph_wall,
        "max_test_rmse_diff_vs_sequential": worst,
        "teams": teams,
        "tune_whole_run_ms_by_solo": tune_solo,
        "tune_whole_run_ms_by_deal": dict(zip(("8", "32", "never"), tune_deal)),
        "problems": [dict(m=bp.m, n=bp.n, nnz=nnz, k=bp.solve.k, updates=bp.solve.outer * bp.solve.k * nnz,
                          sequential_s=t) for bp, (_, _, nnz), t in zip(specs, inits, per_problem)],
    }
    print(json.dumps(out), flush=True)


def bench_foldin(args):
    """Held-out-users fold-in on the 480,189 x 17,770 shape, k = 40: 10 % of
    the users are held out, CCD++ factorizes the rest (10 outer x 5 inner),
    the MLP (82 -> 256 -> 256 -> 40, GELU) is trained on the kept users'
    (features, w_i) pairs and run with bfloat16 weights, and the held-out
    users are folded in from zero and from the warm start.  Both rating
    models of the shape: the independent 1..5 ratings and the planted
    rank-8 ratings (same frequencies)."""
    from cdmf_amd import data as D
    from cdmf_amd import foldin as F

    out = {"metric": "fold-in sweeps / device time to the plain fold-in's test RMSE (warm start vs zero start)",
           "n_gpus": 1, "dtype": "fp32 (MLP inference: bf16 matrix cores, FP32 accumulation)", "studies": {}}
    for name, kind in (("independent ratings", 1), ("planted rank-8 ratings", 3)):
        spec = dict(D.NETFLIX, rating_kind=kind)
        if args.foldin_users:
            scale = args.foldin_users / spec["m"]
            spec.update(m=args.foldin_users, nnz_train=int(spec["nnz_train"] * scale),
                        nnz_test=int(spec["nnz_test"] * scale))
        t = time.perf_counter()
        ds = D.gen_powerlaw(**spec)
        log("%s: generated %dx%d nnz=%d in %.1fs" % (name, ds.m, ds.n, ds.nnz, time.perf_counter() - t))
        cfg = D.FOLDIN_SOLVE
        r = F.experiment(ds, cfg.k, cfg.lam, cfg.outer, cfg.inner, sweeps=8, epochs=20, converge=150,
                         rmse_tol=args.foldin_tol, max_sweeps=args.foldin_max_sweeps, log=log)
        rec = dict(users=ds.m, held_out=len(r.plain_rmse) and None, target_rmse=r.target,
                   plain_sweeps=r.plain_sweeps, warm_sweeps_to_target=r.warm_sweeps,
                   plain_time_to_target_s=r.plain_time[r.plain_sweeps],
                   warm_time_to_target_s=r.warm_time[r.warm_sweeps] if r.warm_sweeps >= 0 else None,
                   plain_rmse=r.plain_rmse, warm_rmse=r.warm_rmse, plain_time_s=r.plain_time, warm_time_s=r.warm_time,
                   objective_star=r.obj_star, plain_objective=r.plain_obj, warm_objective=r.warm_obj,
                   plain_sweeps_to_objective=r.plain_sweeps_obj, warm_sweeps_to_objective=r.warm_sweeps_obj,
                   mlp_train_loss=r.train_loss, timings=r.timings,
                   rmse_star=r.rmse_star, rmse_star_warm=r.rmse_star_warm, rmse_tol=r.rmse_tol,
                   plain_sweeps_to_converged=r.plain_sweeps_conv, warm_sweeps_to_converged=r.warm_sweeps_conv,
                   plain_time_to_converged_s=r.plain_time[r.plain_sweeps_conv] if r.plain_sweeps_conv >= 0 else None,
                   warm_time_to_converged_s=r.warm_time[r.warm_sweeps_conv] if r.warm_sweeps_conv >= 0 else None,
                   converged_by_tolerance={
                       "%g" % tol: dict(plain_sweeps=ps, warm_sweeps=ws,
                                        plain_time_s=r.plain_time[ps] if ps >= 0 else None,
                                        warm_time_s=r.warm_time[ws] if ws >= 0 else None)
                       for tol, (ps, ws) in r.conv_by_tol.items()})
        rec.pop("held_out")
        out["studies"][name] = rec
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
