#!/usr/bin/env python3
"""Benchmark: converged FP64 solves per second of the detector-tomography
problem M=2000, N=1000, D=4000 (alpha^2 on [0,1800], eta=0.75, 1000-pixel
multiplexed read-out, 1e7 shots, gamma=1e-3, eps=1e-7, mode 2), one GPU.

A step is one converged solve from Pi = 1/N with F and P already resident in
HBM. Prints ONE JSON line on stdout; progress goes to stderr.

roofline: the dominant kernel (largest device time over the timed region,
measured with HIP events on the context's stream) with its algorithmic FP64
work per launch, 2 * (stored F elements) * N flops, divided by its average
launch duration.

cpu_baseline: the reference program (oracle/_ref/pqdts_omp.x, built from the
upstream sources) in its own benchmark mode (argument 14 = 1) on the same
inputs, host cores via OpenMP; its per-operator times are combined with this
GPU solve's operator counts into a per-solve CPU time. Without the reference
build, the CPU oracle is timed on a bounded prefix of the same solve instead.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "converged solves/s + p99 latency, batched mixed vs plain FP64, |ΔΠ|∞≤1e-6"
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline_reference(prob, op_counts, threads):
    """Reference bench mode on the same inputs; returns a cpu_baseline dict or None."""
    import oracle
    from pqdts_amd import workloads
    if not os.path.exists(oracle.REF_BIN):
        return None
    with tempfile.TemporaryDirectory(dir="/tmp") as w:
        NF, NP = workloads.write_inputs(w, prob)
        cmd = [oracle.REF_BIN, str(prob.M), str(prob.N), str(prob.D), str(NF), str(NP), "2", "0", "0",
               str(prob.gamma), str(prob.eps), "1", "0", "0", "1"]
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="true", OMP_PLACES="cores",
                   FORT_FMT_RECL="100000")
        t0 = time.time()
        r = subprocess.run(cmd, cwd=w, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=600)
        wall = time.time() - t0
    if r.returncode != 0:
        log("reference bench failed:", r.stdout[-500:])
        return None
    per = {}
    for line in r.stdout.splitlines():
        t = line.split()
        if len(t) >= 3 and t[0] == "t":
            per.setdefault(t[1], []).append(float(t[2]))
    med = {k: float(np.median(v)) for k, v in per.items()}
    # one reference CG iteration = BF + Bdiag (recomputed, pqdts.f90:2065) + 4 pdot + 3 axpy-type sweeps
    n_bf, n_l, n_dl = op_counts.get("BF", 0), op_counts.get("L", 0), op_counts.get("dL", 0)
    n_proj = op_counts.get("doproj", 0)
    t_solve = (n_bf * (med["BF"] + med["Bdiag"] + 4 * med["pdot"] + 2 * med["paxpy"] + med["pgaxpy"])
               + n_l * med["L"] + n_dl * med["dL"] + n_proj * med["doproj"])
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": threads, "kind": "reference",
            "sample": (f"reference pqdts_omp.x benchmark mode (arg 14 = 1) on the same inputs, "
                       f"{threads} OpenMP threads, {wall:.1f} s wall incl. ingest; median per-operator times "
                       f"(BF {med['BF']:.4f} s, dL {med['dL']:.4f} s, L {med['L']:.4f} s, Bdiag {med['Bdiag']:.4f} s) "
                       f"times this solve's counts (BF {n_bf}, L {n_l}, dL {n_dl}, doproj {n_proj}) = "
                       f"{t_solve:.1f} s per solve"),
            "per_op_s": med}


def cpu_baseline_port(prob, budget_s=20.0):
    """The CPU oracle (scalar C) on a bounded prefix of the same solve."""
    import oracle
    o = oracle.Oracle.from_problem(prob)
    o.buf("X")[:] = 1.0 / prob.N
    t0 = time.time()
    _, ev = o.solve(2, 1, prob.eps, 1, 0)   # one Newton iteration per stage
    dt = time.time() - t0
    n_pcg = sum(1 for k, _ in ev if k == oracle.EV_PCG)
    return {"value": None, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"CPU oracle, 1 Newton iteration per stage ({n_pcg} PCG iterations) in {dt:.1f} s",
            "seconds_per_pcg": dt / max(n_pcg, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--maxiter", type=int, default=2000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--kernels", type=int, default=0,
                    help="instead of solving: run each hot operator this many times on the same problem "
                         "(plus a plain 16 MB copy for counter calibration) and exit; for profiler passes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); independent problems shard "
              f"one per GPU with no communication, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import pqdts_amd
    from pqdts_amd import workloads
    if pqdts_amd.device_count() < 1:
        log("no GPU visible")
        sys.exit(1)
    t = time.time()
    prob = workloads.config2()
    log(f"generated cfg2: M={prob.M} N={prob.N} D={prob.D} nnz={prob.nnz} in {time.time() - t:.1f} s")
    ctx = pqdts_amd.Context.from_problem(prob)
    info = ctx.info()
    x0 = 1.0 / prob.N

    if args.kernels:
        rng = np.random.default_rng(0)
        X = rng.random((prob.M, prob.N))
        ctx.put("X", X / X.sum(1, keepdims=True))
        ctx.put("XTMP", rng.normal(size=(prob.M, prob.N)))
        for r in range(args.kernels):
            for s in ("G", "HP", "SN"):  # 48 MB of writes push X out of the 32 MB of L2
                ctx.pset(0.0, s)
            ctx.pcopy("X", "Z")          # 16 MB read + 16 MB write: counter calibration
            ctx.L("X")
            ctx.dL("X", "G")
            ctx.BF("XTMP", "HP")
        log(f"ran {args.kernels} x (pcopy, L, dL, BF) on cfg2")
        return

    def progress(_c, kind, v):
        if kind == 0 and int(v[1]) % 25 == 0:
            log(f"  stage {int(v[0])} iteration {int(v[1])}: L={v[2]:.10e} conv={v[4]:.3e}")

    def one_solve(tag):
        ctx.pset(x0, "X")
        t0 = time.perf_counter()
        conv, ev = ctx.solve(2, args.maxiter, prob.eps, 1, 0, on_event=progress)
        dt = time.perf_counter() - t0
        its = [v for k, v in ev if k == 0]
        npcg = sum(1 for k, _ in ev if k == 1)
        nls = sum(1 for k, _ in ev if k == 2)
        log(f"{tag}: {dt:.3f} s, {len(its)} Newton iterations, {npcg} PCG, {nls} line-search trials, "
            f"L={its[-1][2]:.16e} conv={conv:.3e}")
        return dt, conv, its, npcg, nls

    for w in range(args.warmup):
        one_solve(f"warmup {w}")
    ctx.kernel_timing(True)
    base_timers = ctx.timers()
    lat = []
    t_all = time.perf_counter()
    for s in range(args.steps):
        dt, conv, its, npcg, nls = one_solve(f"step {s}")
        lat.append(dt)
    total = time.perf_counter() - t_all
    kstats = ctx.kernel_stats()
    timers = ctx.timers()
    op_counts = {k: (timers[k][0] - base_timers.get(k, (0, 0))[0]) // args.steps for k in timers}

    # roofline of the dominant kernel
    dom = max(kstats, key=lambda k: kstats[k][1])
    launches, ms = kstats[dom]
    avg_s = ms / 1e3 / launches
    # algorithmic FP64 work per launch: stored elements of the banded operand x N columns x 2
    work = {"spmm_F": 2.0 * info["nspan"] * prob.N, "spmmT_F": 2.0 * info["nspan"] * prob.N,
            "hvp_G": 2.0 * info["gram_nnz"] * prob.N}
    flops = work[dom]
    achieved = flops / avg_s / 1e12
    roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": dom,
            "avg_launch_us": avg_s * 1e6, "launches_per_step": launches / args.steps,
            "flops_per_launch": flops,
            "kernels": {k: {"launches": v[0], "total_ms": v[1], "avg_us": 1e3 * v[1] / max(v[0], 1)}
                        for k, v in kstats.items()}}

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        try:
            cpu = cpu_baseline_reference(prob, op_counts, threads)
            if cpu is None:
                cpu = cpu_baseline_port(prob)
        except Exception as e:  # the baseline is reported, never fatal
            log("cpu baseline failed:", repr(e))

    lat_sorted = np.sort(np.array(lat))
    out = {
        "metric": METRIC,
        "value": args.steps / total,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "cfg2 plain FP64 solve, one problem at a time", "M": prob.M, "N": prob.N,
                   "D": prob.D, "nnzF": prob.nnz, "gamma": prob.gamma, "eps": prob.eps, "mode": 2,
                   "eta": 0.75, "shots": 1e7, "npix": 1000,
                   "hvp": "gram" if info["gram"] else "F products"},
        "latency_s": {"p50": float(np.percentile(lat_sorted, 50)), "p99": float(np.percentile(lat_sorted, 99)),
                      "max": float(lat_sorted[-1])},
        "solve": {"newton_iterations": len(its), "pcg_iterations": npcg, "line_search_trials": nls,
                  "final_L": its[-1][2], "final_conv": conv, "converged": bool(conv < prob.eps),
                  "op_counts": op_counts},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
