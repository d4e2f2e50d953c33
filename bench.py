#!/usr/bin/env python3
"""Benchmark of the detector-tomography solver on one MI355X.

Metric: converged solves/s (+ p99 latency), batched mixed precision vs plain FP64,
with ||dPi||_inf <= 1e-6 between the two. Prints ONE JSON line on stdout;
progress goes to stderr.

Workloads (--workload):
  mix     (default) 256 independent problems of uneven size: M log-uniform in
          [50, 2000], N uniform in [M/4, M], D = 2M, eta in [0.5, 0.95], gamma
          log-uniform in [1e-5, 1e-2], 1e6-1e8 shots, solved to eps (default 1e-7,
          the mid-size case's tolerance).
          A step = the whole mix solved by the batched engine (one launch per
          operator over all unfinished problems), mixed precision. `value` is its
          converged solves/s. Beside it: the same engine in FP64, and the plain
          FP64 solver run one problem at a time (pqdts_solve) on every
          --seq-stride-th problem of the mix, with ||dPi||_inf and objective gaps
          of the batched-mixed results against those plain FP64 results, beside the
          same gaps between two FP64 solves that differ only in summation order.
  cfg2    one problem M=2000, N=1000, D=4000 (alpha^2 on [0,1800], eta=0.75,
          1000-pixel read-out, 1e7 shots, gamma=1e-3, eps=1e-7): plain FP64 solves/s.
  served  Poisson arrivals (50 req/s) of mix-distributed requests plus 10% large
          ones (M=8000, N=4000, D=16000), admitted by HBM footprint into the
          batched engine as they arrive; p50/p99 latency from arrival to result.

Inputs are resident in HBM before timing starts (except in `served`, where the
upload is part of serving a request). roofline: the dominant kernel over the
timed region, its algorithmic work (2 x stored elements of its banded operand x
N per problem and launch) over its HIP-event time on the engine's stream.
cpu_baseline: the reference program (oracle/_ref/pqdts_omp.x, built from the
upstream sources) solving a bounded sample of the same problems on the host
cores, or, without that build, the CPU oracle on a bounded sample.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "converged solves/s + p99 latency, batched mixed vs plain FP64, |ΔΠ|∞≤1e-6"
FP64_PEAK = 78.6    # TFLOP/s, MI355X FP64 (vector and matrix)
FP32_PEAK = 157.3   # TFLOP/s, MI355X FP32 matrix


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(x, q):
    return float(np.percentile(np.asarray(x, dtype=np.float64), q)) if len(x) else None


HBM_PEAK = 8000.0    # GB/s, MI355X HBM3E


def roofline_of(kstats):
    """Dominant kernel by device time. Band products are priced against the MFMA
    peak of their precision (achieved = algorithmic flops / time), streaming
    kernels against HBM (achieved = algorithmic bytes / time)."""
    if not kstats:
        return None
    dom = max(kstats, key=lambda k: kstats[k][1])
    launches, ms, flops, nbytes = kstats[dom]
    sec = ms / 1e3
    per = {k: {"launches": v[0], "total_ms": v[1], "avg_us": 1e3 * v[1] / max(v[0], 1),
               "tflops": (v[2] / (v[1] / 1e3) / 1e12) if v[1] > 0 else 0.0,
               "gbps": (v[3] / (v[1] / 1e3) / 1e9) if v[1] > 0 else 0.0} for k, v in kstats.items()}
    if flops > 0:
        peak = FP32_PEAK if dom.endswith("32") else FP64_PEAK
        achieved = flops / sec / 1e12 if sec > 0 else 0.0
        out = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s"}
    else:
        achieved = nbytes / sec / 1e9 if sec > 0 else 0.0
        out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK, "unit": "GB/s"}
    out.update({"frac": out["achieved"] / out["peak"], "traffic": None, "kernel": dom,
                "avg_launch_us": 1e3 * ms / max(launches, 1), "launches": launches,
                "flops_per_launch": flops / max(launches, 1), "bytes_per_launch": nbytes / max(launches, 1),
                "kernels": per})
    return out


# ------------------------------------------------------------------ CPU baseline
def ref_solve_seconds(prob, eps, maxiter, threads, timeout):
    """Solves one problem with the reference binary; returns (wall s, final X) or None."""
    import oracle
    from pqdts_amd import fortio, workloads
    with tempfile.TemporaryDirectory(dir="/tmp") as w:
        NF, NP = workloads.write_inputs(w, prob)
        cmd = [oracle.REF_BIN, str(prob.M), str(prob.N), str(prob.D), str(NF), str(NP), "2", str(maxiter), "1",
               str(prob.gamma), str(eps), "1", "0", "0", "0"]
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="true", OMP_PLACES="cores")
        t0 = time.time()
        try:
            subprocess.run(cmd, cwd=w, env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                           timeout=timeout, check=True)
        except (subprocess.TimeoutExpired, subprocess.CalledProcessError):
            return None
        dt = time.time() - t0
        _, _, _, _, _, X = fortio.read_state(os.path.join(w, "rank_     0_final.dat"))
    return dt, X


def ref_bench_mode(prob, threads, timeout=300):
    """The reference binary's benchmark mode (argument 14 = 1, pqdts.f90:1581-1734): median
    seconds per operator call ("t <op>" lines), and the wall time of the run."""
    import oracle
    from pqdts_amd import workloads
    with tempfile.TemporaryDirectory(dir="/tmp") as w:
        NF, NP = workloads.write_inputs(w, prob)
        cmd = [oracle.REF_BIN, str(prob.M), str(prob.N), str(prob.D), str(NF), str(NP), "2", "0", "0",
               str(prob.gamma), str(prob.eps), "1", "0", "0", "1"]
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="true", OMP_PLACES="cores",
                   FORT_FMT_RECL="100000")
        t0 = time.time()
        r = subprocess.run(cmd, cwd=w, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=timeout)
        wall = time.time() - t0
    per = {}
    for line in r.stdout.splitlines():
        t = line.split()
        if len(t) >= 3 and t[0] == "t":
            per.setdefault(t[1], []).append(float(t[2]))
    return {k: float(np.median(v)) for k, v in per.items()}, wall


def solve_seconds_from_ops(med, op_counts):
    """A reference solve's time from its per-operator times and a solve's operator counts.
    One reference CG step = BF + Bdiag (recomputed every step, pqdts.f90:2065) + 4 pdot +
    2 paxpy + pgaxpy; the line search and the iteration evaluate L, dL and the projection."""
    n_bf, n_l, n_dl = op_counts.get("BF", 0), op_counts.get("L", 0), op_counts.get("dL", 0)
    n_proj = op_counts.get("doproj", 0) + op_counts.get("doproj2", 0)
    return (n_bf * (med["BF"] + med["Bdiag"] + 4 * med["pdot"] + 2 * med["paxpy"] + med["pgaxpy"])
            + n_l * med["L"] + n_dl * med["dL"] + n_proj * med["doproj"])


def cpu_baseline_mix(problems, order, mid, eps, maxiter, threads, budget_s, gpu_seconds, mid_counts):
    """The reference binary on the host cores: full solves of small problems of the sample (in
    `order`) until budget_s is spent, plus one mid-size problem (`mid`) whose solve time is its
    benchmark-mode operator times x the operator counts of the same problem's plain GPU solve."""
    import oracle
    if not os.path.exists(oracle.REF_BIN):
        return cpu_baseline_port(problems[order[0]])
    done, spent, names, gpu_t = 0, 0.0, [], 0.0
    for i in order:
        if spent > budget_s:
            break
        r = ref_solve_seconds(problems[i], eps, maxiter, threads, timeout=max(5.0, 2 * budget_s - spent))
        if r is None:
            break
        spent += r[0]
        done += 1
        names.append(f"M={problems[i].M},N={problems[i].N}")
        gpu_t += gpu_seconds.get(i, float("nan"))
    mid_txt, t_mid = "", 0.0
    if mid is not None:
        med, wall = ref_bench_mode(problems[mid], threads)
        t_mid = solve_seconds_from_ops(med, mid_counts)
        gpu_t += gpu_seconds.get(mid, float("nan"))
        mid_txt = (f"; plus M={problems[mid].M},N={problems[mid].N} estimated at {t_mid:.1f} s from the reference's "
                   f"benchmark mode (arg 14 = 1, {wall:.1f} s run: BF {med['BF'] * 1e3:.2f} ms, L {med['L'] * 1e3:.2f} "
                   f"ms, dL {med['dL'] * 1e3:.2f} ms per call) x the operator counts of its plain GPU solve "
                   f"({', '.join(f'{k} {v}' for k, v in sorted(mid_counts.items()) if k in ('BF', 'L', 'dL', 'doproj', 'doproj2'))})")
    n = done + (1 if mid is not None else 0)
    if n == 0:
        return None
    total = spent + t_mid
    return {"value": n / total, "unit": "solves/s", "cores": threads, "kind": "reference",
            "sample": (f"reference pqdts_omp.x ({threads} OpenMP threads, built from the upstream sources) solving "
                       f"{done} problems of the mix ({', '.join(names)}) to eps={eps:g}: {spent:.1f} s{mid_txt}; "
                       f"the plain FP64 GPU solver takes {gpu_t:.2f} s for the same problems"),
            "gpu_plain_fp64_same_sample_solves_per_s": n / gpu_t if gpu_t > 0 else None}


def cpu_baseline_port(prob):
    import oracle
    o = oracle.Oracle.from_problem(prob)
    o.buf("X")[:] = 1.0 / prob.N
    t0 = time.time()
    _, ev = o.solve(2, 1, prob.eps, 1, 0)
    dt = time.time() - t0
    n_pcg = sum(1 for k, _ in ev if k == oracle.EV_PCG)
    return {"value": None, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"CPU oracle, 1 Newton iteration per stage of M={prob.M} ({n_pcg} PCG) in {dt:.1f} s"}


def cpu_baseline_cfg2(prob, op_counts, threads):
    """Reference bench mode (arg 14 = 1) per-operator times x this solve's operator counts."""
    import oracle
    if not os.path.exists(oracle.REF_BIN):
        return cpu_baseline_port(prob)
    med, wall = ref_bench_mode(prob, threads, timeout=600)
    t_solve = solve_seconds_from_ops(med, op_counts)
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": threads, "kind": "reference",
            "sample": (f"reference pqdts_omp.x benchmark mode (arg 14 = 1) on the same inputs, {threads} OpenMP "
                       f"threads, {wall:.1f} s wall incl. ingest; median per-operator times (BF {med['BF']:.4f} s, "
                       f"dL {med['dL']:.4f} s, L {med['L']:.4f} s, Bdiag {med['Bdiag']:.4f} s) x this solve's "
                       f"counts (BF {op_counts.get('BF', 0)}, L {op_counts.get('L', 0)}, dL {op_counts.get('dL', 0)}) "
                       f"= {t_solve:.1f} s per solve")}


# ------------------------------------------------------------------ workloads
HVP_SPLIT = -1  # --hvp-split: -1 leaves the library's choice
POLISH = 0      # --polish: outer iterations of the polish after conv < eps (0: none)
POLISH_CLASSES = 2  # --polish-classes
POLISH_BUDGET = 1.0  # --polish-budget: polish CG rounds, at most this multiple of the solve's PCG rounds
POLISH_METHOD = "direct"  # --polish-method: "direct" face solves or "cg" (Newton-CG steps)


def make_ctx(pqdts_amd, prob, polish=True):
    c = pqdts_amd.Context.from_problem(prob)
    if HVP_SPLIT >= 0:
        c.set_hvp_split(HVP_SPLIT)
    if POLISH > 0 and polish:
        c.set_polish(POLISH, budget=POLISH_BUDGET, fp32=True, method=POLISH_METHOD)
    return c


def polish_set(problems, classes):
    """Indices of the problems in the `classes` smallest of the lanes' three size classes
    (thirds by M x N): the ones the polish is applied to."""
    order = sorted(range(len(problems)), key=lambda i: problems[i].M * problems[i].N)
    return set(order[:len(order) * max(0, min(3, classes)) // 3])


def lane_groups(problems, lanes):
    order = sorted(range(len(problems)), key=lambda i: problems[i].M * problems[i].N)
    k = max(1, min(lanes, len(problems)))
    return [order[len(order) * g // k: len(order) * (g + 1) // k] for g in range(k)]


def run_batched(pqdts_amd, problems, ctxs, precision, maxiter, timing, tag, lanes=1, lane_cus=None):
    """All problems from Pi = 1/N through the batched engine; returns (seconds, latencies, X list,
    results, kernel stats). With lanes > 1 the problems are split by size (M x N, equal counts) into
    that many batches, each on its own stream and driven by its own host thread, so that small
    problems advance at their own pace instead of at the pace of the largest ones. lane_cus (one
    entry per lane, smallest problems first: a range of CU indices, or None for all) restricts
    where each lane's workgroups run. Every batch is created, restricted and timed in this thread
    before any lane thread starts. timing: per-kernel HIP events on every lane's stream."""
    import threading
    for p, c in zip(problems, ctxs):
        c.pset(1.0 / p.N, "X")
    groups = [g for g in lane_groups(problems, lanes) if g]
    batches = []
    for g, idx in enumerate(groups):
        b = pqdts_amd.Batch(len(idx))
        if lane_cus and len(groups) > 1 and lane_cus[g] is not None:
            b.set_cus(lane_cus[g])
        if timing:
            b.kernel_timing(timing)
        batches.append(b)
    lat, res = [None] * len(problems), [None] * len(problems)
    kst, errors = {}, []
    kst_lane = [None] * len(groups)
    lock = threading.Lock()
    t0 = time.perf_counter()

    def lane(g, idx, b):
        try:
            slot_of, steps = {}, 0
            for i in idx:
                slot_of[b.admit(ctxs[i], 2, maxiter, problems[i].eps, 1, 0, precision, record=False)] = i
            t_log = time.perf_counter()
            while slot_of:
                for s in b.step():
                    i = slot_of.pop(s)
                    lat[i] = time.perf_counter() - t0
                    res[i] = b.result(s)
                    b.release(s)
                    res[i]["polish"] = ctxs[i].polish_result()
                steps += 1
                if steps % 1000 == 0 or time.perf_counter() - t_log > 30.0:
                    t_log = time.perf_counter()
                    log(f"  {tag} lane {g}: {steps} steps, {len(slot_of)} unfinished, "
                        f"{time.perf_counter() - t0:.1f} s")
            log(f"  {tag} lane {g}: {len(idx)} problems done in {steps} steps, {time.perf_counter() - t0:.1f} s")
            if timing:
                mine = b.kernel_stats()
                kst_lane[g] = mine
                with lock:
                    for name, v in mine.items():
                        old = kst.get(name, (0, 0.0, 0.0, 0.0))
                        kst[name] = tuple(x + y for x, y in zip(old, v))
        except Exception as e:  # reported by the caller
            errors.append(e)

    threads = [threading.Thread(target=lane, args=(g, idx, b)) for g, (idx, b) in enumerate(zip(groups, batches))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    dt = time.perf_counter() - t0
    for b in batches:
        b.close()
    if errors:
        raise errors[0]
    X = [c.get("X") for c in ctxs]
    if timing:
        kst["__lanes__"] = kst_lane
    return dt, lat, X, res, kst


def lane_rooflines(kst):
    """The dominant band product of each lane (smallest problems first): launches, average
    duration, achieved TFLOP/s, fraction of its MFMA peak."""
    out = []
    for ks in kst.get("__lanes__") or []:
        if not ks:
            out.append(None)
            continue
        prods = {k: v for k, v in ks.items() if v[2] > 0}
        if not prods:
            out.append(None)
            continue
        k = max(prods, key=lambda n: prods[n][1])
        n, ms, fl, _ = prods[k]
        peak = FP32_PEAK if k.endswith("32") else FP64_PEAK
        tf = fl / (ms / 1e3) / 1e12 if ms else 0.0
        out.append({"kernel": k, "launches": n, "avg_us": 1e3 * ms / max(n, 1), "tflops": tf, "frac": tf / peak})
    return out


def objective(ctx):
    return ctx.L("X")


def calibration_copy(pqdts_amd, problems, ctxs):
    """Copies of a known size (Z = X of the largest problem, 5 times) for scaling
    the FETCH_SIZE / WRITE_SIZE counters of a profiler pass."""
    i = max(range(len(problems)), key=lambda k: problems[k].M * problems[k].N)
    c = ctxs[i]
    for _ in range(5):
        c.pcopy("X", "Z")
    c.get("Z")
    nbytes = 8 * problems[i].M * problems[i].N
    log(f"calibration copy (the last GPU work of this run): 5 x kb_axpby reading and writing {nbytes} bytes "
        f"each (M={problems[i].M}, N={problems[i].N})")
    return nbytes


def parse_lane_cus(args, pqdts_amd):
    """--lane-cus: "none" (every lane on every CU); "auto" (3 lanes: the smallest problems on
    the first eighth of the CUs alone, the middle lane on all, the largest problems on all but
    the first quarter); counts "a,b,c" (consecutive exclusive partitions); or one consecutive
    range per lane, "lo-hi" or "all" (e.g. 0-31,all,64-255). Returns one range (or None for
    all CUs) per lane."""
    spec = args.lane_cus.strip()
    ncu = pqdts_amd.device_cu_count()
    if spec in ("", "none"):
        return None
    if spec == "auto":
        if args.lanes != 3 or ncu < 16:
            return None
        return [range(0, ncu // 8), None, range(ncu // 4, ncu)]
    items = spec.split(",")
    if len(items) != args.lanes:
        raise SystemExit(f"--lane-cus needs one entry per lane ({args.lanes})")
    if all(x.isdigit() for x in items):
        cus = [int(x) for x in items]
        if sum(cus) > ncu or min(cus) < 1:
            raise SystemExit("--lane-cus counts must be positive and sum to at most the device's CUs")
        return [range(sum(cus[:g]), sum(cus[:g + 1])) for g in range(len(cus))]
    out = []
    for x in items:
        if x == "all":
            out.append(None)
            continue
        parts = x.split("-")
        if len(parts) != 2 or not all(v.isdigit() for v in parts):
            raise SystemExit(f"--lane-cus entry {x}: expected lo-hi or all")
        lo, hi = int(parts[0]), int(parts[1])
        if not 0 <= lo <= hi < ncu:
            raise SystemExit(f"--lane-cus range {x} outside 0-{ncu - 1}")
        out.append(range(lo, hi + 1))
    return out


def lane_cus_str(lane_cus):
    if lane_cus is None:
        return None
    return ["all" if r is None else f"{r.start}-{r.stop - 1}" for r in lane_cus]


PMC_SUMMARY = os.path.join(ROOT, "profiles", "mix256_hvp_G32_traffic.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed two-pass rocprofv3 PMC summary
    (FETCH_SIZE, WRITE_SIZE, each scaled by a copy kernel of known size), or None. The
    summary's machine-readable line is `traffic <kernel> <bytes per launch>`."""
    try:
        for line in open(PMC_SUMMARY):
            t = line.strip().strip("`").split()
            if len(t) == 3 and t[0] == "traffic" and t[1] == kernel:
                return float(t[2])
    except OSError:
        pass
    return None


def determined_subset(args, pqdts_amd, problems, ctxs, eps=1e-11):
    """The problems of the mix whose optimum FP64 pins (M <= 260, gamma >= 3e-4) solved
    to eps = 1e-11: batched mixed precision against the plain FP64 solver one problem at
    a time, with the 1e-6 gate on every problem."""
    import dataclasses
    idx = [i for i, p in enumerate(problems) if p.M <= 260 and p.gamma >= 3e-4]
    if not idx:
        return None
    sub = [dataclasses.replace(problems[i], eps=eps) for i in idx]
    sctx = [ctxs[i] for i in idx]
    dt_m, lat_m, X_m, res_m, _ = run_batched(pqdts_amd, sub, sctx, 1, args.maxiter, False, "determined, mixed")
    t_plain, X_p, conv_p = 0.0, [], []
    for p, c in zip(sub, sctx):
        c.pset(1.0 / p.N, "X")
        t0 = time.perf_counter()
        conv, _ = c.solve(2, args.maxiter, eps, 1, 0)
        t_plain += time.perf_counter() - t0
        X_p.append(c.get("X"))
        conv_p.append(conv < eps)
    dpi = [float(np.max(np.abs(a - b))) for a, b in zip(X_m, X_p)]
    gated = sum(1 for r, d, cp in zip(res_m, dpi, conv_p) if r["conv"] < eps and cp and d <= 1e-6)
    rate_m, rate_p = gated / dt_m, sum(conv_p) / t_plain
    log(f"determined subset: {len(sub)} problems at eps {eps:g}: mixed {dt_m:.2f} s, {gated} within 1e-6; "
        f"plain fp64 {t_plain:.2f} s")
    return {"problems": len(sub), "selection": "M <= 260 and gamma >= 3e-4", "eps": eps,
            "mixed_within_1e-6": gated, "mixed_seconds": dt_m, "mixed_solves_per_s_gated": rate_m,
            "plain_fp64_seconds": t_plain, "plain_fp64_solves_per_s": rate_p,
            "speedup_vs_plain_fp64": rate_m / rate_p if rate_p > 0 else None,
            "dpi_inf_max": max(dpi), "dpi_inf_median": pct(dpi, 50), "p99_latency_s": pct(lat_m, 99)}


def certified(r):
    """The polish (if on) certified the result as the problem's minimiser."""
    return POLISH == 0 or (r.get("polish") or {}).get("status") == 1


def polish_summary(res_m, res_b64, problems):
    if POLISH == 0:
        return None
    out = {"method": POLISH_METHOD, "max_outer": POLISH, "budget": POLISH_BUDGET,
           "size_classes_polished": POLISH_CLASSES,
           "not_polished": sum(1 for r in res_m if r["polish"]["status"] == 0),
           "stopped_uncertified": [{"M": problems[i].M, "N": problems[i].N, "gamma": problems[i].gamma,
                                    "mixed": res_m[i]["polish"], "fp64": res_b64[i]["polish"]}
                                   for i in range(len(problems))
                                   if 2 in (res_m[i]["polish"]["status"], res_b64[i]["polish"]["status"])][:12]}
    order = sorted(range(len(problems)), key=lambda i: problems[i].M * problems[i].N)
    thirds = [order[len(order) * g // 3: len(order) * (g + 1) // 3] for g in range(3)]
    out["certified_by_size_third"] = {
        tag: [sum(1 for i in t if res[i]["polish"]["status"] == 1) for t in thirds]
        for tag, res in (("mixed", res_m), ("fp64", res_b64))}
    out["cg_rounds_by_size_third"] = {
        tag: [int(sum(res[i]["polish"]["cg_rounds"] for i in t)) for t in thirds]
        for tag, res in (("mixed", res_m), ("fp64", res_b64))}
    for tag, res in (("mixed", res_m), ("fp64", res_b64)):
        pr = [r["polish"] for r in res]
        out[tag] = {"certified": sum(1 for q in pr if q["status"] == 1), "of": len(pr),
                    "iterations": int(sum(q["iterations"] for q in pr)),
                    "max_iterations": int(max(q["iterations"] for q in pr)),
                    "cg_rounds": int(sum(q["cg_rounds"] for q in pr)),
                    "max_cg_rounds": int(max(q["cg_rounds"] for q in pr))}
    return out


def workload_mix(args, pqdts_amd, workloads):
    T_START = time.time()
    problems = workloads.batch_mix(args.count, eps=args.eps)
    log(f"generated {len(problems)} problems in {time.time() - T_START:.1f} s; M median "
        f"{int(np.median([p.M for p in problems]))}, max {max(p.M for p in problems)}")
    if pqdts_amd.device_count() < 1:
        raise SystemExit("no GPU visible")
    t = time.time()
    pol_idx = polish_set(problems, args.polish_classes)
    ctxs = [make_ctx(pqdts_amd, p, i in pol_idx) for i, p in enumerate(problems)]
    n_gram = sum(1 for c in ctxs if c.info()["gram"])
    log(f"uploaded in {time.time() - t:.1f} s; {n_gram} problems take the Gram Hessian product")
    lane_cus = parse_lane_cus(args, pqdts_amd)
    # warmup: the 8 smallest problems, a few Newton iterations each, at most warmup rounds
    small = sorted(range(len(problems)), key=lambda i: problems[i].M * problems[i].N)[:8]
    tw = time.time()
    for w in range(args.warmup):
        if time.time() - tw > 15.0:
            break
        run_batched(pqdts_amd, [problems[i] for i in small], [ctxs[i] for i in small], 1, 20, False,
                    f"warmup {w}", args.lanes, lane_cus)
    # timed steps: the whole mix solved by the batched engine in mixed precision, with per-kernel
    # HIP events on every lane's stream (the roofline source); as many of the requested steps as
    # fit the wall-clock budget (at least one)
    runs, kst_m = [], {}
    t_timed = time.perf_counter()
    for s in range(args.steps):
        dt, lat, Xm, res, kst = run_batched(pqdts_amd, problems, ctxs, 1, args.maxiter, args.ksample,
                                            f"batched mixed step {s}", args.lanes, lane_cus)
        lanes_kst = kst.pop("__lanes__", None)
        for name, v in kst.items():
            old = kst_m.get(name, (0, 0.0, 0.0, 0.0))
            kst_m[name] = tuple(x + y for x, y in zip(old, v))
        if s == 0:
            lanes_rf = lane_rooflines({"__lanes__": lanes_kst})
        if runs and not all(np.array_equal(a, b) for a, b in zip(runs[0][2], Xm)):
            raise SystemExit("batched mixed results differ between steps")
        runs.append((dt, lat, Xm if not runs else runs[0][2], res))
        log(f"step {s}: batched mixed {dt:.2f} s")
        spent = time.perf_counter() - t_timed
        if spent + dt > args.budget:
            break
    steps = len(runs)
    total = sum(r[0] for r in runs)
    dt_m, lat_m, X_m, res_m = runs[0]
    if args.legs == "timed":
        # only the timed region (for profiler runs): the line carries the timing and the roofline
        rf = roofline_of(kst_m)
        if rf is not None:
            rf["per_lane"] = lanes_rf
        return {"metric": METRIC, "value": None, "unit": "solves/s", "n_gpus": 1, "steps": steps,
                "warmup": args.warmup, "ms_per_step": 1e3 * total / steps, "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "f32 Hessian products / f64 elsewhere",
                "data": "synthetic", "config": {"workload": "mix timed region only (--legs timed)",
                                                "count": len(problems), "eps": args.eps},
                "converged_solves_per_s_ungated": steps * sum(r["conv"] < p.eps for r, p in zip(res_m, problems))
                / total, "roofline": rf, "cpu_baseline": None,
                "calibration_copy_bytes": calibration_copy(pqdts_amd, problems, ctxs) if args.calib_copy else None}
    Lm = []
    for c, x in zip(ctxs, X_m):
        c.put("X", x)
        Lm.append(objective(c))
    # the FP64 result of every problem: the batched engine in FP64, which is bit-identical to the
    # plain solver run alone (checked below on the plain sample)
    dt_b64, lat_b64, X_b64, res_b64, _ = run_batched(pqdts_amd, problems, ctxs, 0, args.maxiter, False,
                                                     "batched fp64", args.lanes, lane_cus)
    L64 = []
    for c, x in zip(ctxs, X_b64):
        c.put("X", x)
        L64.append(objective(c))
    # converged: the reference's test conv < eps; with the polish on, also certified
    # as the problem's minimiser by the polish
    conv_m = [r["conv"] < p.eps for r, p in zip(res_m, problems)]
    conv_64 = [r["conv"] < p.eps for r, p in zip(res_b64, problems)]
    cert_m = [c and certified(r) for c, r in zip(conv_m, res_m)]
    cert_64 = [c and certified(r) for c, r in zip(conv_64, res_b64)]
    dpi = [float(np.max(np.abs(a - b))) for a, b in zip(X_m, X_b64)]
    gap = [abs(a - b) / abs(b) for a, b in zip(Lm, L64)]
    gated = [cm and c64 and d <= 1e-6 for cm, c64, d in zip(cert_m, cert_64, dpi)]
    log(f"batched fp64: {dt_b64:.2f} s; mixed within 1e-6 of fp64: {sum(gated)}/{len(problems)}")
    # plain FP64, one problem at a time (pqdts_solve), on every seq_stride-th problem
    seq_idx = list(range(0, len(problems), args.seq_stride))
    mid = max((i for i in seq_idx if problems[i].M <= 1000), key=lambda i: problems[i].M, default=None)
    seq_t, conv_p, cert_p, same_bits, mid_counts = {}, [], [], 0, {}
    for n, i in enumerate(seq_idx):
        p, c = problems[i], ctxs[i]
        c.pset(1.0 / p.N, "X")
        base = c.timers()
        t0 = time.perf_counter()
        conv, _ = c.solve(2, args.maxiter, p.eps, 1, 0)
        seq_t[i] = time.perf_counter() - t0
        if i == mid:
            after = c.timers()
            mid_counts = {k: after[k][0] - base.get(k, (0, 0))[0] for k in after}
        conv_p.append(conv < p.eps)
        cert_p.append(conv < p.eps and certified({"polish": c.polish_result()}))
        same_bits += int(np.array_equal(c.get("X"), X_b64[i]))
        if n % 8 == 0 or POLISH > 0:
            log(f"  plain fp64 {n + 1}/{len(seq_idx)}: M={p.M} N={p.N} {seq_t[i]:.3f} s")
    seq_total = sum(seq_t.values())
    plain_rate = sum(conv_p) / seq_total
    plain_cert_rate = sum(cert_p) / seq_total
    # FP64 determinacy floor: the plain sample in FP64 with the other Hessian-product summation
    # order (Gram <-> F products), under the same 1e-6 gate
    sample = [problems[i] for i in seq_idx]
    sctx = [ctxs[i] for i in seq_idx]
    for c in sctx:
        c.set_hvp_path("F" if c.info()["gram"] else "G")
    _, _, X_alt, _, _ = run_batched(pqdts_amd, sample, sctx, 0, args.maxiter, False,
                                    "fp64, other summation order", args.lanes, lane_cus)
    for c in sctx:
        c.set_hvp_path("auto")
    floor = [float(np.max(np.abs(X_b64[i] - x))) for i, x in zip(seq_idx, X_alt)]
    det = determined_subset(args, pqdts_amd, problems, ctxs) if POLISH == 0 else None
    nconv_m = sum(conv_m)
    value = steps * sum(gated) / total
    seq_lat = np.cumsum([seq_t[i] for i in seq_idx]) * args.seq_stride
    rf = roofline_of(kst_m)
    if rf is not None:
        rf["per_lane"] = lanes_rf
        tr = measured_traffic(rf["kernel"])
        rf["traffic"] = tr
        rf["traffic_source"] = os.path.relpath(PMC_SUMMARY, ROOT) if tr is not None else None
    out = {
        "metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * total / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 Hessian products / f64 elsewhere", "data": "synthetic",
        "config": {"workload": f"mix of {len(problems)} uneven problems (M log-uniform [50,2000], N U[M/4,M], "
                               f"D=2M, alpha^2 on [0,0.9M], eta U[0.5,0.95], gamma log-U[1e-5,1e-2], shots "
                               f"1e6-1e8), eps={args.eps:g}, mode 2, from Pi=1/N, batched engine in {args.lanes} "
                               f"size lanes (one stream and host thread each), mixed precision",
                   "count": len(problems), "eps": args.eps, "maxiter_per_stage": args.maxiter,
                   "lanes": args.lanes, "lane_cus": lane_cus_str(lane_cus)},
        "steps_requested": args.steps, "timed_budget_s": args.budget,
        "value_definition": "batched mixed solves per second that converged (FP64 KKT test conv < eps), "
                            + ("were certified as the problem's minimiser by the polish (as was the FP64 "
                               "result), " if POLISH > 0 else "")
                            + "and are within ||dPi||inf <= 1e-6 of the same problem's plain FP64 result",
        "converged": {"batched_mixed": nconv_m, "of": len(problems), "batched_fp64": int(sum(conv_64)),
                      "mixed_within_1e-6_of_fp64": int(sum(gated)),
                      "certified_mixed": int(sum(cert_m)), "certified_fp64": int(sum(cert_64)),
                      "plain_fp64_sample": int(sum(conv_p)), "plain_fp64_sample_certified": int(sum(cert_p)),
                      "plain_sample_size": len(seq_idx)},
        "converged_solves_per_s_ungated": steps * nconv_m / total,
        "latency_s": {"p50": pct(lat_m, 50), "p99": pct(lat_m, 99), "max": max(lat_m),
                      "note": "every problem submitted at t=0"},
        "plain_fp64_one_at_a_time": {
            "solves_per_s": plain_rate, "certified_solves_per_s": plain_cert_rate,
            "sample": f"every {args.seq_stride}th problem of the mix ({len(seq_idx)} problems), solved alone with "
                      f"pqdts_solve; the whole mix one at a time would take {seq_total * args.seq_stride:.1f} s",
            "seconds": seq_total, "bit_identical_to_batched_fp64": f"{same_bits}/{len(seq_idx)}",
            "latency_s_fifo_extrapolated": {"p50": pct(seq_lat, 50), "p99": pct(seq_lat, 99)}},
        # like for like: gated mixed solves against plain FP64 solves under the same
        # convergence definition (certified when the polish is on)
        "speedup_vs_plain_fp64": value / (plain_cert_rate if POLISH > 0 else plain_rate)
        if (plain_cert_rate if POLISH > 0 else plain_rate) > 0 else None,
        "speedup_vs_plain_fp64_ungated": (steps * nconv_m / total) / plain_rate if plain_rate > 0 else None,
        "batched_fp64": {"solves_per_s": sum(conv_64) / dt_b64, "seconds": dt_b64, "p50": pct(lat_b64, 50),
                         "p99": pct(lat_b64, 99)},
        "accuracy_mixed_vs_fp64": {
            "problems_compared": len(dpi), "dpi_inf_max": max(dpi), "dpi_inf_median": pct(dpi, 50),
            "dpi_le_1e-6": int(sum(d <= 1e-6 for d in dpi)),
            "rel_objective_gap_max": max(gap), "rel_objective_gap_median": pct(gap, 50),
            "fp64_vs_fp64_sample": len(floor), "fp64_vs_fp64_dpi_median": pct(floor, 50),
            "fp64_vs_fp64_le_1e-6": int(sum(f <= 1e-6 for f in floor)),
            "note": "fp64_vs_fp64 = ||dPi||inf between two FP64 solves of the same problem whose Hessian products "
                    "sum in a different order (Gram vs F products): how far eps leaves Pi undetermined; the 1e-6 "
                    "gate cannot hold where FP64 itself does not"},
        "determined_subset": det,
        "polish": polish_summary(res_m, res_b64, problems),
        "slowest_mixed": [{"M": problems[i].M, "N": problems[i].N, "gamma": problems[i].gamma,
                           "newton": res_m[i]["newton"], "pcg": res_m[i]["pcg"], "latency_s": lat_m[i],
                           "polish": res_m[i].get("polish")}
                          for i in sorted(range(len(problems)), key=lambda k: -lat_m[k])[:5]],
        "solve_stats_mixed": {"newton": int(sum(r["newton"] for r in res_m)),
                              "pcg": int(sum(r["pcg"] for r in res_m)),
                              "max_newton": int(max(r["newton"] for r in res_m))},
        "roofline": rf,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        rng = np.random.default_rng(7)
        order = [i for i in rng.permutation(seq_idx) if problems[i].M <= 400]
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        try:
            out["cpu_baseline"] = cpu_baseline_mix(problems, order, mid, args.eps, args.maxiter, threads,
                                                   args.cpu_budget, seq_t, mid_counts)
        except Exception as e:
            log("cpu baseline failed:", repr(e))
    if args.calib_copy:
        out["calibration_copy_bytes"] = calibration_copy(pqdts_amd, problems, ctxs)
    out["bench_wall_s"] = time.time() - T_START
    return out


def workload_cfg2(args, pqdts_amd, workloads):
    t = time.time()
    prob = workloads.config2()
    log(f"generated cfg2: M={prob.M} N={prob.N} D={prob.D} nnz={prob.nnz} in {time.time() - t:.1f} s")
    if pqdts_amd.device_count() < 1:
        raise SystemExit("no GPU visible")
    ctx = make_ctx(pqdts_amd, prob)
    info = ctx.info()

    def progress(_c, kind, v):
        if kind == 0 and int(v[1]) % 25 == 0:
            log(f"  stage {int(v[0])} iteration {int(v[1])}: L={v[2]:.10e} conv={v[4]:.3e}")

    def one(tag):
        ctx.pset(1.0 / prob.N, "X")
        t0 = time.perf_counter()
        conv, ev = ctx.solve(2, args.maxiter, prob.eps, 1, 0, on_event=progress)
        dt = time.perf_counter() - t0
        its = [v for k, v in ev if k == 0]
        npcg = sum(1 for k, _ in ev if k == 1)
        nls = sum(1 for k, _ in ev if k == 2)
        log(f"{tag}: {dt:.3f} s, {len(its)} Newton, {npcg} PCG, {nls} line-search trials, "
            f"L={its[-1][2]:.16e} conv={conv:.3e}")
        return dt, conv, its, npcg, nls

    for w in range(args.warmup):
        one(f"warmup {w}")
    ctx.kernel_timing(True)
    base = ctx.timers()
    lat = []
    t_all = time.perf_counter()
    for s in range(args.steps):
        dt, conv, its, npcg, nls = one(f"step {s}")
        lat.append(dt)
    total = time.perf_counter() - t_all
    kst = ctx.kernel_stats()
    timers = ctx.timers()
    counts = {k: (timers[k][0] - base.get(k, (0, 0))[0]) // args.steps for k in timers}
    out = {
        "metric": METRIC, "value": args.steps / total, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * total / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "cfg2 plain FP64 solve, one problem at a time", "M": prob.M, "N": prob.N,
                   "D": prob.D, "nnzF": prob.nnz, "gamma": prob.gamma, "eps": prob.eps, "mode": 2,
                   "hvp": "gram" if info["gram"] else "F products", "hvp_split": info["hvp_split"]},
        "latency_s": {"p50": pct(lat, 50), "p99": pct(lat, 99), "max": max(lat)},
        "solve": {"newton_iterations": len(its), "pcg_iterations": npcg, "line_search_trials": nls,
                  "final_L": its[-1][2], "final_conv": conv, "converged": bool(conv < prob.eps), "op_counts": counts},
        "roofline": roofline_of(kst),
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline_cfg2(prob, counts, args.cpu_threads or min(16, os.cpu_count() or 1))
        except Exception as e:
            log("cpu baseline failed:", repr(e))
    return out


def workload_cfg3(args, pqdts_amd, workloads):
    """The single large instance: FP64 and mixed precision, each a solve capped at
    --maxiter Newton iterations per stage from Pi = 1/N, every kernel timed."""
    t = time.time()
    prob = workloads.config3()
    log(f"generated cfg3: M={prob.M} N={prob.N} D={prob.D} nnz={prob.nnz} in {time.time() - t:.1f} s")
    if pqdts_amd.device_count() < 1:
        raise SystemExit("no GPU visible")
    t = time.time()
    ctx = make_ctx(pqdts_amd, prob)
    t_up = time.time() - t
    info = ctx.info()
    log(f"context in {t_up:.1f} s: {info}")
    runs = {}
    for prec in (0, 1):
        ctx.pset(1.0 / prob.N, "X")
        b = pqdts_amd.Batch(1)
        b.kernel_timing(True)
        b.admit(ctx, 2, args.maxiter, prob.eps, 1, 0, prec, record=True)
        t0 = time.perf_counter()
        while b.active():
            fin = b.step()
            its = [v for k, v in b.events(0) if k == 0]
            if its:
                log(f"  {'mixed' if prec else 'fp64'}: stage {int(its[-1][0])} iteration {int(its[-1][1])} "
                    f"L={its[-1][2]:.10e} conv={its[-1][4]:.3e} {time.perf_counter() - t0:.1f} s")
            if fin:
                break
        dt = time.perf_counter() - t0
        res = b.result(0)
        kst = b.kernel_stats()
        its = [v for k, v in b.events(0) if k == 0]
        b.close()
        x = ctx.get("X")
        runs[prec] = {"seconds": dt, "res": res, "kst": kst, "X": x, "L": objective(ctx), "iters": len(its)}
        log(f"{'mixed' if prec else 'fp64'}: {dt:.2f} s, {res['newton']} Newton, {res['pcg']} PCG, "
            f"L={runs[prec]['L']:.12e}")
    r64, r32 = runs[0], runs[1]
    dpi = float(np.max(np.abs(r64["X"] - r32["X"])))
    gap = abs(r32["L"] - r64["L"]) / abs(r64["L"])
    same_active = bool(np.array_equal(r64["X"] < 1e-8, r32["X"] < 1e-8))
    rate = r32["res"]["newton"] / r32["seconds"]
    return {
        "metric": METRIC, "value": rate, "unit": "Newton iterations/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": 1e3 * r32["seconds"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 Hessian products / f64 elsewhere", "data": "synthetic",
        "config": {"workload": f"single large instance M={prob.M}, N={prob.N}, D={prob.D} (alpha^2 on [0,18000], "
                               f"eta=0.9, saturation at N-1, 1e8 shots, gamma=1e-2), mode 2 from Pi=1/N, capped at "
                               f"{args.maxiter} Newton iterations per stage", "nnzF": prob.nnz,
                   "hvp": "gram" if info["gram"] else "F products", "device_bytes": info.get("device_bytes")},
        "fp64": {"seconds": r64["seconds"], "newton": r64["res"]["newton"], "pcg": r64["res"]["pcg"],
                 "L": r64["L"], "s_per_newton": r64["seconds"] / max(r64["res"]["newton"], 1),
                 "roofline": roofline_of(r64["kst"])},
        "mixed": {"seconds": r32["seconds"], "newton": r32["res"]["newton"], "pcg": r32["res"]["pcg"],
                  "L": r32["L"], "s_per_newton": r32["seconds"] / max(r32["res"]["newton"], 1)},
        "accuracy_mixed_vs_fp64": {"dpi_inf": dpi, "rel_objective_gap": gap, "same_active_set_1e-8": same_active,
                                   "note": "both capped at the same Newton iteration counts per stage"},
        "upload_s": t_up,
        "roofline": roofline_of(r32["kst"]),
        "cpu_baseline": None,
    }


def workload_served(args, pqdts_amd, workloads):
    """Poisson arrivals of mix-distributed requests plus 10 % large ones, served by the
    library's request server (pqdts_server_*): three size-class lanes with continuous
    batching, admission against a device-byte budget; latency = result time - arrival,
    upload included. Beside it: the plain FP64 solver serving the same stream one request
    at a time in arrival order."""
    from pqdts_amd import _lib
    rng = np.random.default_rng(11)
    n = args.count
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, n))
    small = workloads.batch_mix(n, seed=5, eps=args.eps)
    large_idx = sorted(int(i) for i in rng.choice(n, size=max(1, n // 10), replace=False))
    specs = [(8000, 4000, 16000, 0.8, 1e-3, 1e8, 900 + i, args.eps, f"large{i}") for i in large_idx]
    larges = dict(zip(large_idx, workloads.build_many(specs)))
    reqs = [larges.get(i, small[i]) for i in range(n)]
    coos = [_lib.coo_of(p) for p in reqs]
    log(f"served: {n} requests over {arrivals[-1]:.1f} s, {len(larges)} large")
    if pqdts_amd.device_count() < 1:
        raise SystemExit("no GPU visible")
    budget = int(args.hbm_gb * 2**30)
    bounds = (20000, 400000)  # M x N of the lane boundaries
    # the mix's lane layout ("auto") slows the large requests that bound this stream:
    # every lane shares every CU unless ranges are given
    lane_cus = None if args.lane_cus.strip() == "auto" else parse_lane_cus(args, pqdts_amd)

    def polished_req(i):
        """The polish applies to the requests of the size classes below the largest one
        (as in the mix bench): the largest problems are not certified within its budget."""
        return reqs[i].M * reqs[i].N < bounds[min(len(bounds), POLISH_CLASSES) - 1] if POLISH_CLASSES < 3 else True

    def serve(precision):
        srv = pqdts_amd.Server(bounds, lane_cus=lane_cus, capacity=args.capacity or 256, budget_bytes=budget)
        lat, X, conv, cert, ids = [None] * n, [None] * n, [None] * n, [None] * n, {}
        pol_now = None
        t0 = time.perf_counter()
        nxt = 0
        while nxt < n or len(ids):
            now = time.perf_counter() - t0
            while nxt < n and arrivals[nxt] <= now:
                if POLISH > 0 and polished_req(nxt) != pol_now:
                    # the server applies its polish settings to the requests submitted after them
                    pol_now = polished_req(nxt)
                    srv.set_polish(POLISH if pol_now else 0, budget=POLISH_BUDGET, fp32=True, method=POLISH_METHOD)
                ids[srv.submit(reqs[nxt], maxiter=args.maxiter, precision=precision, coo=coos[nxt])] = nxt
                nxt += 1
            wait_ms = int(max(1.0, 1e3 * (arrivals[nxt] - now))) if nxt < n else 1000
            for rid in srv.wait(timeout_ms=min(wait_ms, 1000)):
                i = ids.pop(rid)
                r = srv.result(rid)
                lat[i], X[i], conv[i] = r["latency_s"], r["X"], r["conv"]
                cert[i] = certified(r)
        dt = time.perf_counter() - t0
        info = srv.info()
        srv.close()
        return dt, lat, X, conv, cert, info

    def serve_plain():
        """Plain FP64 solver, one request at a time in arrival order (upload included)."""
        lat, X, cert = [None] * n, [None] * n, [None] * n
        t0 = time.perf_counter()
        for i in range(n):
            now = time.perf_counter() - t0
            if arrivals[i] > now:
                time.sleep(arrivals[i] - now)
            (fr, fc, fv), (pr, pc, pv) = coos[i]
            c = pqdts_amd.Context(reqs[i].M, reqs[i].N, reqs[i].D, (fr, fc, fv), (pr, pc, pv), reqs[i].gamma)
            if POLISH > 0 and polished_req(i):
                c.set_polish(POLISH, budget=POLISH_BUDGET, method=POLISH_METHOD)
            c.pset(1.0 / reqs[i].N, "X")
            c.solve(2, args.maxiter, reqs[i].eps, 1, 0)
            X[i] = c.get("X")
            cert[i] = certified({"polish": c.polish_result()})
            lat[i] = time.perf_counter() - t0 - arrivals[i]
            c.close()
            if i % 16 == 0:
                log(f"  plain fp64 request {i + 1}/{n}: {time.perf_counter() - t0:.1f} s")
        return time.perf_counter() - t0, lat, X, cert

    dt_m, lat_m, X_m, conv_m, cert_m, info_m = serve(1)
    log(f"served mixed: {dt_m:.1f} s, p50 {pct(lat_m, 50):.2f} s, p99 {pct(lat_m, 99):.2f} s, "
        f"peak {info_m['peak_bytes'] / 2**30:.2f} GiB of {budget / 2**30:.1f}")
    dt_p, lat_p, X_p, cert_p = serve_plain()
    log(f"served plain fp64 one at a time: {dt_p:.1f} s, p50 {pct(lat_p, 50):.2f} s, p99 {pct(lat_p, 99):.2f} s")
    dpi = [float(np.max(np.abs(a - b))) for a, b in zip(X_m, X_p)]
    gated = sum(1 for d, c, p, cm, cp in zip(dpi, conv_m, reqs, cert_m, cert_p)
                if c < p.eps and cm and cp and d <= 1e-6)
    lat_large = [lat_m[i] for i in large_idx]
    return {
        "metric": METRIC, "value": gated / dt_m, "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": 1e3 * dt_m, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 Hessian products / f64 elsewhere", "data": "synthetic",
        "config": {"workload": f"served stream: {n} requests, Poisson arrivals at {args.rate:g}/s from the mix "
                               f"distribution, 10% large (M=8000,N=4000,D=16000), eps={args.eps:g}, library request "
                               f"server: 3 size-class lanes, admission by device footprint "
                               f"({args.hbm_gb:g} GiB budget), upload included"},
        "value_definition": "served mixed solves per second that converged, "
                            + ("were certified as the problem's minimiser by the polish (as was the FP64 "
                               "result), " if POLISH > 0 else "")
                            + "and are within ||dPi||inf <= 1e-6 of the plain FP64 result",
        "polish": {"method": POLISH_METHOD, "max_outer": POLISH, "budget": POLISH_BUDGET,
                   "certified_mixed": int(sum(map(bool, cert_m))),
                   "certified_fp64": int(sum(map(bool, cert_p)))} if POLISH > 0 else None,
        "converged_solves_per_s_ungated": sum(1 for c, p in zip(conv_m, reqs) if c < p.eps) / dt_m,
        "within_1e-6": gated,
        "latency_s": {"p50": pct(lat_m, 50), "p99": pct(lat_m, 99), "max": max(lat_m),
                      "large_p50": pct(lat_large, 50), "large_max": max(lat_large)},
        "admission": {"budget_bytes": budget, "peak_bytes": info_m["peak_bytes"]},
        "plain_fp64_one_at_a_time": {"solves_per_s": n / dt_p, "p50": pct(lat_p, 50), "p99": pct(lat_p, 99)},
        "speedup_vs_plain_fp64_ungated": dt_p / dt_m,
        "accuracy_mixed_vs_fp64": {"dpi_inf_median": pct(dpi, 50), "dpi_le_1e-6": int(sum(d <= 1e-6 for d in dpi))},
        "roofline": None, "cpu_baseline": None,
    }


def kernels_probe(args, pqdts_amd, workloads):
    """Each hot operator a few times on cfg2 plus a 16 MB copy, for profiler passes."""
    prob = workloads.config2()
    ctx = make_ctx(pqdts_amd, prob)
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["mix", "cfg2", "cfg3", "served"], default="mix")
    ap.add_argument("--count", type=int, default=0, help="problems (mix: 256) or requests (served: 64)")
    ap.add_argument("--eps", type=float, default=1e-7)
    ap.add_argument("--maxiter", type=int, default=5000)
    ap.add_argument("--capacity", type=int, default=0)
    ap.add_argument("--lanes", type=int, default=3)
    ap.add_argument("--lane-cus", default="auto",
                    help="CUs per lane, smallest problems first: auto, none, counts (48,80,128) or ranges "
                         "(0-31,all,64-255)")
    ap.add_argument("--seq-stride", type=int, default=8)
    ap.add_argument("--rate", type=float, default=50.0)
    ap.add_argument("--hbm-gb", type=float, default=200.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--legs", choices=["all", "timed"], default="all",
                    help="mix: all measurements, or only the warmup and timed steps (profiler runs)")
    ap.add_argument("--ksample", type=int, default=16,
                    help="per-kernel HIP-event timing in the timed steps: every k-th operator launch")
    ap.add_argument("--budget", type=float, default=100.0,
                    help="wall-clock seconds for the timed steps (mix): steps run while they fit")
    ap.add_argument("--kernels", type=int, default=0)
    ap.add_argument("--calib-copy", action="store_true")
    ap.add_argument("--polish", type=int, default=60,
                    help="polish every solve to the problem's minimiser after conv < eps, at most this many "
                         "primal-dual active-set iterations (0: off)")
    ap.add_argument("--polish-classes", type=int, default=2,
                    help="polish the problems of this many of the three size classes (thirds by M x N), "
                         "smallest first; measured: the largest third is not certified within the budget")
    ap.add_argument("--polish-budget", type=float, default=1.0,
                    help="polish CG rounds, at most this multiple of the minimisation's own PCG rounds "
                         "(the Newton-CG polish)")
    ap.add_argument("--polish-method", choices=("direct", "cg"), default="direct",
                    help="the polish's face solves: direct (column-block inverses and the row sums' band "
                         "Schur complement) or Newton-CG steps")
    ap.add_argument("--hvp-split", type=int, default=-1,
                    help="split of the Gram product over band chunks (0 by shape, 1 none, k fixed)")
    args = ap.parse_args()
    global HVP_SPLIT, POLISH, POLISH_BUDGET, POLISH_CLASSES, POLISH_METHOD
    HVP_SPLIT = args.hvp_split
    POLISH_METHOD = args.polish_method
    # the polish applies to the mix (the metric's 1e-6 gate); the other workloads report
    # the reference's own stopping point
    POLISH, POLISH_BUDGET = (args.polish if args.workload in ("mix", "served") else 0), args.polish_budget
    POLISH_CLASSES = args.polish_classes
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); independent problems would "
              f"shard one per GPU with no communication, so N GPUs serve N times the problems of one",
              file=sys.stderr)
        sys.exit(2)
    import pqdts_amd
    from pqdts_amd import workloads
    if args.kernels:
        kernels_probe(args, pqdts_amd, workloads)
        return
    if args.workload == "cfg2":
        if args.maxiter == 5000:
            args.maxiter = 2000
        out = workload_cfg2(args, pqdts_amd, workloads)
    elif args.workload == "cfg3":
        if args.maxiter == 5000:
            args.maxiter = 3
        out = workload_cfg3(args, pqdts_amd, workloads)
    elif args.workload == "served":
        args.count = args.count or 64
        out = workload_served(args, pqdts_amd, workloads)
    else:
        args.count = args.count or 256
        out = workload_mix(args, pqdts_amd, workloads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
