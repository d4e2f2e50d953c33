"""Summaries of rocprofv3 rocpd databases (kernel trace and PMC passes).

  python -m pqdts_amd.profsum trace <db>                 per-kernel calls / avg / total time
  python -m pqdts_amd.profsum pmc <db> [<db> ...]         per-kernel mean counter value per dispatch
                                                          (+ VGPRs, LDS bytes, grid)
  python -m pqdts_amd.profsum last <db> <kernel> <k>      mean counter values of the last k dispatches
                                                          of the kernel whose name starts with <kernel>
  python -m pqdts_amd.profsum streams <db> <kernel>       dispatches of the kernels whose name starts
                                                          with <kernel>, per stream: calls / avg / total
  python -m pqdts_amd.profsum lanes <db> [<window_s>]     per stream: dispatches, kernel time, span, the
                                                          gaps between its kernels; then each stream's
                                                          busy share in windows of <window_s> seconds
Kernel names are shortened to their template head.
"""
import sqlite3
import sys
from collections import defaultdict


def short(name: str, keep_args: bool = False) -> str:
    n = name.split("(")[0] if not keep_args else name
    return n.replace("void ", "").replace("pq::", "")


def trace(db: str):
    c = sqlite3.connect(db)
    rows = c.execute("select name, count(*), avg(duration), sum(duration), min(duration), max(duration) "
                     "from kernels group by name order by sum(duration) desc").fetchall()
    total = sum(r[3] for r in rows)
    print(f"{'kernel':58s} {'calls':>7s} {'avg_us':>10s} {'total_ms':>10s} {'pct':>6s}")
    for name, calls, avg, tot, mn, mx in rows:
        print(f"{short(name)[:58]:58s} {calls:7d} {avg / 1e3:10.2f} {tot / 1e6:10.2f} {100 * tot / total:6.2f}")
    print(f"{'total':58s} {'':7s} {'':10s} {total / 1e6:10.2f}")


def pmc(dbs):
    agg = defaultdict(lambda: defaultdict(list))
    meta = {}
    for db in dbs:
        c = sqlite3.connect(db)
        for kname, cname, val, vg, avg_, lds, gs, dur in c.execute(
                "select kernel_name, counter_name, value, vgpr_count, accum_vgpr_count, lds_block_size, "
                "grid_size, duration from counters_collection"):
            agg[short(kname)][cname].append(val)
            meta[short(kname)] = (vg, avg_, lds, gs)
    counters = sorted({cn for k in agg for cn in agg[k]})
    print(f"{'kernel':58s} {'dispatches':>10s} " + " ".join(f"{cn + '/disp':>16s}" for cn in counters)
          + f" {'vgpr':>5s} {'agpr':>5s} {'lds_B':>7s} {'grid':>9s}")
    for k in sorted(agg):
        n = max(len(v) for v in agg[k].values())
        vals = [sum(agg[k][cn]) / len(agg[k][cn]) if agg[k][cn] else float("nan") for cn in counters]
        vg, ag, lds, gs = meta[k]
        print(f"{k[:58]:58s} {n:10d} " + " ".join(f"{v:16.1f}" for v in vals) + f" {vg:5d} {ag:5d} {lds:7d} {gs:9d}")


def last(db: str, kernel: str, k: int):
    c = sqlite3.connect(db)
    rows = c.execute("select dispatch_id, kernel_name, counter_name, value, grid_size from counters_collection "
                     "order by dispatch_id").fetchall()
    per = defaultdict(dict)
    for did, kname, cname, val, gs in rows:
        if short(kname).startswith(kernel):
            per[did][cname] = val
            per[did]["grid"] = gs
    ids = sorted(per)[-k:]
    counters = sorted({cn for d in ids for cn in per[d] if cn != "grid"})
    for cn in counters:
        vals = [per[d][cn] for d in ids if cn in per[d]]
        print(f"{kernel} last {len(ids)} dispatches (grid {per[ids[-1]]['grid']}): {cn} mean "
              f"{sum(vals) / len(vals):.1f} per dispatch, values {', '.join(f'{v:.0f}' for v in vals)}")


def streams(db: str, kernel: str):
    c = sqlite3.connect(db)
    cols = [r[1] for r in c.execute("pragma table_info(kernels)").fetchall()]
    key = next((k for k in ("stream_id", "queue_id", "stream", "queue") if k in cols), None)
    if key is None:
        raise SystemExit(f"no stream or queue column in kernels: {cols}")
    rows = c.execute(f"select name, {key}, count(*), avg(duration), sum(duration) from kernels "
                     f"group by name, {key} order by name, min(start)").fetchall()
    print(f"{'kernel':58s} {key:>10s} {'calls':>7s} {'avg_us':>10s} {'total_ms':>10s}")
    for name, sid, calls, avg, tot in rows:
        if short(name).startswith(kernel):
            print(f"{short(name)[:58]:58s} {str(sid):>10s} {calls:7d} {avg / 1e3:10.2f} {tot / 1e6:10.2f}")


def _stream_key(c):
    cols = [r[1] for r in c.execute("pragma table_info(kernels)").fetchall()]
    key = next((k for k in ("stream_id", "queue_id", "stream", "queue") if k in cols), None)
    if key is None:
        raise SystemExit(f"no stream or queue column in kernels: {cols}")
    return key


def lanes(db: str, window_s: float = 5.0):
    """Where each stream's time goes: its kernels (busy), the span from its first
    kernel's start to its last kernel's end, and the gaps between consecutive kernels
    (time the stream waited: on the host, or on its own launches' latency). Kernels of
    one stream do not overlap, so busy = summed duration."""
    c = sqlite3.connect(db)
    key = _stream_key(c)
    rows = c.execute(f"select {key}, start, duration from kernels order by {key}, start").fetchall()
    per = defaultdict(list)
    for sid, start, dur in rows:
        per[sid].append((start, start + dur))
    t0 = min(iv[0][0] for iv in per.values())
    t1 = max(iv[-1][1] for iv in per.values())
    print(f"{key:>10s} {'kernels':>8s} {'busy_ms':>10s} {'span_ms':>10s} {'busy%':>6s} {'gaps_ms':>10s} "
          f"{'gaps>1ms':>8s} {'max_gap_ms':>10s}")
    for sid in sorted(per, key=str):
        iv = per[sid]
        busy = sum(e - s_ for s_, e in iv)
        span = iv[-1][1] - iv[0][0]
        gaps = [b[0] - a[1] for a, b in zip(iv, iv[1:]) if b[0] > a[1]]
        print(f"{str(sid):>10s} {len(iv):8d} {busy / 1e6:10.2f} {span / 1e6:10.2f} {100 * busy / span:6.1f} "
              f"{sum(gaps) / 1e6:10.2f} {sum(1 for g in gaps if g > 1e6):8d} {max(gaps, default=0) / 1e6:10.2f}")
    w = int(window_s * 1e9)
    nw = max(1, -(-(t1 - t0) // w))
    print(f"busy share per {window_s:g} s window (from the first kernel):")
    print(f"{'window_s':>9s} " + " ".join(f"{str(sid):>8s}" for sid in sorted(per, key=str)))
    for k in range(nw):
        a, b = t0 + k * w, t0 + (k + 1) * w
        shares = []
        for sid in sorted(per, key=str):
            busy = sum(max(0, min(e, b) - max(s_, a)) for s_, e in per[sid])
            shares.append(busy / w)
        print(f"{k * window_s:9.3f} " + " ".join(f"{x:8.2f}" for x in shares))


if __name__ == "__main__":
    if sys.argv[1] == "lanes":
        lanes(sys.argv[2], float(sys.argv[3]) if len(sys.argv) > 3 else 5.0)
    elif sys.argv[1] == "streams":
        streams(sys.argv[2], sys.argv[3])
    elif sys.argv[1] == "trace":
        trace(sys.argv[2])
    elif sys.argv[1] == "last":
        last(sys.argv[2], sys.argv[3], int(sys.argv[4]))
    else:
        pmc(sys.argv[2:])
