"""Offline autotuner of the Gram Hessian product's split over band chunks.

The Hessian product 2 G p runs one workgroup per 64 x 64 output tile over the
tile's band chunks. A problem with fewer tiles than the GPU holds at once runs at
the latency of its longest tile; splitting each tile's chunks over k workgroups
(k partial arrays, then a summing pass) shortens that chain k-fold at the price of
the partials' traffic. Which k wins depends on the number of output tiles and on
the chunks of the longest tile, and on the machine.

This tuner runs on the GPU. For a grid of problem shapes spanning the uneven
benchmark mix (M 50..2000, N from M/4 to M, D = 2M) and beyond (M up to 8000), it
builds each problem's Gram band, puts it in the reduced space the CG works in,
and times the FP64 Hessian product (pqdts_BF, every launch between HIP events)
with the split k = 1, 2 and 4. Per bucket of (log2 output tiles, log2 most chunks
of a tile) it keeps the fastest k (k = 1 unless a split is at least 3 % faster) and
writes pqdts_amd/csrc/tuned_split.inc, which the library's split_by_shape reads.
Buckets no tuned shape reached keep the library's rule.

The table is committed with the measurements behind it: a problem's split, and
so its bits, depends on its shape and the library build only, never on when or
where it runs, so a problem computes the same bits alone and in any batch.

    python -m pqdts_amd.autotune [--reps 12] [--out pqdts_amd/csrc/tuned_split.inc]
"""
from __future__ import annotations

import argparse
import datetime
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TILE_BUCKETS, CHUNK_BUCKETS = 11, 10
SPLITS = (1, 2, 4)
GAIN = 0.03


def shapes():
    for M in (50, 80, 130, 200, 320, 500, 800, 1300, 2000, 3200, 5000, 8000):
        for frac in (0.25, 0.5, 1.0):
            yield M, max(2, int(M * frac)), 2 * M


def bucket(v: int, n: int) -> int:
    return min(n - 1, max(0, int(v).bit_length() - 1))


def time_bf(ctx, reps: int) -> float:
    """Microseconds of device time per Hessian product (every kernel of pqdts_BF)."""
    best = float("inf")
    for _ in range(3):
        ctx.kernel_timing(1)
        for _ in range(reps):
            ctx.BF("XTMP", "HP")
        ks = ctx.kernel_stats()
        ctx.kernel_timing(0)
        best = min(best, 1e3 * sum(v[1] for v in ks.values()) / reps)
    return best


def measure(M, N, D, reps, rng):
    import pqdts_amd
    from pqdts_amd import workloads
    F = workloads.poisson_rows(np.linspace(0.0, 0.9 * M, D), M).tocoo()
    empty = (np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0))
    ctx = pqdts_amd.Context(M, N, D, (F.row, F.col, F.data), empty, 1e-3)
    ctx.set_hvp_path("G")
    X = rng.random((M, N)) ** 3
    ctx.put("X", X / X.sum(1, keepdims=True))
    ctx.fullspace = False
    ctx.transform()
    ctx.put("XTMP", rng.normal(size=(M, N)))
    info = ctx.info()
    ntile = -(-N // 64) * -(-M // 64)
    out = {}
    for k in SPLITS:
        ctx.set_hvp_split(k)
        ctx.BF("XTMP", "HP")  # warm
        out[k] = time_bf(ctx, reps)
    ctx.close()
    return ntile, info["gram_nch"], out


def choose(times: dict) -> int:
    best = min(times, key=times.get)
    return best if times[best] < (1.0 - GAIN) * times[1] else 1


def write_table(path, rows, device):
    table = [[0] * CHUNK_BUCKETS for _ in range(TILE_BUCKETS)]
    votes = {}
    for M, N, D, ntile, nch, times in rows:
        key = (bucket(ntile, TILE_BUCKETS), bucket(nch, CHUNK_BUCKETS))
        votes.setdefault(key, []).append(times)
    for (b1, b2), ts in votes.items():
        tot = {k: sum(t[k] for t in ts) for k in SPLITS}
        table[b1][b2] = choose(tot)
    lines = [
        "// Generated by python -m pqdts_amd.autotune: the fastest measured split of the FP64 Gram",
        "// Hessian product over band chunks per bucket [log2 output tiles][log2 most chunks of a",
        "// tile]; 0 = no tuned shape in the bucket (split_by_shape falls back to its rule).",
        f"// Device: {device}; {datetime.date.today().isoformat()}. Measured (us per Hessian product,"
        " reduced space, split 1 / 2 / 4):",
    ]
    for M, N, D, ntile, nch, times in rows:
        lines.append(f"//   M={M:5d} N={N:5d} D={D:5d} tiles={ntile:6d} chunks={nch:4d}: "
                     + " / ".join(f"{times[k]:9.1f}" for k in SPLITS) + f" -> {choose(times)}")
    lines.append(f"constexpr int TUNED_TILE_BUCKETS = {TILE_BUCKETS};")
    lines.append(f"constexpr int TUNED_CHUNK_BUCKETS = {CHUNK_BUCKETS};")
    lines.append("static const unsigned char TUNED_SPLIT[TUNED_TILE_BUCKETS][TUNED_CHUNK_BUCKETS] = {")
    for b1 in range(TILE_BUCKETS):
        lines.append("    {" + ", ".join(str(v) for v in table[b1]) + "},")
    lines.append("};")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=12)
    ap.add_argument("--out", default=os.path.join(HERE, "csrc", "tuned_split.inc"))
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    import pqdts_amd
    if pqdts_amd.device_count() < 1:
        raise SystemExit("the autotuner needs a GPU")
    rng = np.random.default_rng(0)
    rows = []
    t0 = time.time()
    for M, N, D in shapes():
        ntile, nch, times = measure(M, N, D, args.reps, rng)
        rows.append((M, N, D, ntile, nch, times))
        print(f"M={M} N={N} tiles={ntile} chunks={nch}: " + " ".join(f"k={k} {times[k]:.1f} us" for k in SPLITS)
              + f" -> {choose(times)}  ({time.time() - t0:.0f} s)", flush=True)
    write_table(args.out, rows, f"gfx950, {pqdts_amd.device_cu_count()} CUs")
    print("wrote", args.out)


if __name__ == "__main__":
    main()
