"""Text summaries of rocprofv3 results databases (rocpd SQLite).

    python -m vtkmakecompare_amd.profsummary trace  <results.db> [--steps K]
    python -m vtkmakecompare_amd.profsummary traffic <fetch.db> <write.db> \
        --points N [--ref KERNEL --ref-bytes B]

`trace` lists per-kernel call counts and average/total durations of a
--kernel-trace run; `traffic` lists FETCH_SIZE and WRITE_SIZE per dispatch
(KiB as reported, and in units of one interior double field of N points),
optionally scaled by what a reference kernel of known traffic reads.
"""
from __future__ import annotations

import argparse
import collections
import sqlite3


def _short(name: str) -> str:
    return name.split("(")[0].replace("void ", "")


def trace(db: str, steps: int | None = None) -> str:
    c = sqlite3.connect(db)
    rows = c.execute("select name, count(*), avg(duration), sum(duration) from kernels "
                     "group by name order by sum(duration) desc").fetchall()
    total = sum(r[3] for r in rows)
    out = [f"{'kernel':70s} {'calls':>6s} {'avg_us':>10s} {'total_ms':>10s} {'pct':>6s}"]
    for name, n, avg, tot in rows:
        out.append(f"{_short(name)[:70]:70s} {n:6d} {avg / 1e3:10.2f} {tot / 1e6:10.3f} "
                   f"{100.0 * tot / total:6.2f}")
    if steps:
        out.append(f"total kernel time {total / 1e6:.3f} ms over the run")
    return "\n".join(out)


def _per_kernel(db: str, counter: str):
    c = sqlite3.connect(db)
    d = collections.defaultdict(list)
    for name, val in c.execute("select kernel_name, value from counters_collection "
                               "where counter_name = ?", (counter,)):
        d[_short(name)].append(val)
    return {k: (sum(v) / len(v), len(v)) for k, v in d.items()}


def traffic(fetch_db: str, write_db: str, points: int, ref: str | None = None,
            ref_bytes: float | None = None) -> str:
    f = _per_kernel(fetch_db, "FETCH_SIZE")
    w = _per_kernel(write_db, "WRITE_SIZE")
    field = points * 8.0
    scale_f = scale_w = 1.0
    out = []
    if ref and ref in f and ref_bytes:
        scale_f = ref_bytes / (f[ref][0] * 1024.0)
        scale_w = ref_bytes / (w[ref][0] * 1024.0) if ref in w else 1.0
        out.append(f"scaled by reference kernel {ref}: fetch x{scale_f:.3f}, write x{scale_w:.3f}")
    out.append(f"{'kernel':70s} {'fetch_MB':>10s} {'fields':>7s} {'write_MB':>10s} {'fields':>7s} "
               f"{'n':>5s}")
    for k in sorted(f, key=lambda k: -f[k][0]):
        fb = f[k][0] * 1024.0 * scale_f
        wb = w.get(k, (0.0, 0))[0] * 1024.0 * scale_w
        out.append(f"{k[:70]:70s} {fb / 1e6:10.1f} {fb / field:7.2f} {wb / 1e6:10.1f} "
                   f"{wb / field:7.2f} {f[k][1]:5d}")
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("db")
    t.add_argument("--steps", type=int)
    m = sub.add_parser("traffic")
    m.add_argument("fetch_db")
    m.add_argument("write_db")
    m.add_argument("--points", type=int, required=True)
    m.add_argument("--ref")
    m.add_argument("--ref-bytes", type=float)
    a = ap.parse_args()
    if a.cmd == "trace":
        print(trace(a.db, a.steps))
    else:
        print(traffic(a.fetch_db, a.write_db, a.points, a.ref, a.ref_bytes))


if __name__ == "__main__":
    main()
