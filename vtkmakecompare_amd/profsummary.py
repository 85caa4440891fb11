"""Text summaries of rocprofv3 results databases (rocpd SQLite).

    python -m vtkmakecompare_amd.profsummary trace  <results.db> [--steps K]
    python -m vtkmakecompare_amd.profsummary traffic <fetch.db> <write.db> \
        --points N [--calib-fetch F.db --calib-write W.db --calib-bytes B]
    python -m vtkmakecompare_amd.profsummary counters <pmc.db> COUNTER... \
        [--trace-db T.db] [--match NAME]

`trace` lists per-kernel call counts and average/total durations of a
--kernel-trace run; `traffic` lists FETCH_SIZE and WRITE_SIZE per dispatch
scaled to bytes by what the same counters read for a copy of known size
(vtkmakecompare_amd.calibrate), and in units of one interior double field
of N points.
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


def calib_scale(fetch_db: str, write_db: str, kernel: str, nbytes: float):
    """Factors that turn FETCH_SIZE and WRITE_SIZE (KiB per dispatch) into
    bytes, from a copy kernel that reads and writes `nbytes` per launch."""
    f = _per_kernel(fetch_db, "FETCH_SIZE")
    w = _per_kernel(write_db, "WRITE_SIZE")
    if kernel not in f or kernel not in w:
        raise SystemExit(f"calibration kernel {kernel} not found")
    return nbytes / (f[kernel][0] * 1024.0), nbytes / (w[kernel][0] * 1024.0)


def traffic(fetch_db: str, write_db: str, points: int, fetch_scale: float = 1.0,
            write_scale: float = 1.0) -> str:
    """Per-kernel memory traffic per dispatch; `fields` is the traffic in
    units of one interior double field of `points` points."""
    f = _per_kernel(fetch_db, "FETCH_SIZE")
    w = _per_kernel(write_db, "WRITE_SIZE")
    field = points * 8.0
    out = [f"counter scale: fetch x{fetch_scale:.4f}, write x{write_scale:.4f}",
           f"{'kernel':64s} {'fetch_MB':>10s} {'fields':>7s} {'write_MB':>10s} {'fields':>7s} "
           f"{'total_B':>14s} {'n':>5s}"]
    for k in sorted(f, key=lambda k: -f[k][0]):
        fb = f[k][0] * 1024.0 * fetch_scale
        wb = w.get(k, (0.0, 0))[0] * 1024.0 * write_scale
        out.append(f"{k.replace(' ', '')[:64]:64s} {fb / 1e6:10.1f} {fb / field:7.2f} "
                   f"{wb / 1e6:10.1f} {wb / field:7.2f} {fb + wb:14.0f} {f[k][1]:5d}")
    return "\n".join(out)


def _durations(db: str):
    """Average duration (ns) and count per kernel of a --kernel-trace run."""
    c = sqlite3.connect(db)
    return {_short(n): (avg, cnt) for n, cnt, avg in
            c.execute("select name, count(*), avg(duration) from kernels group by name")}


FP64_COUNTERS = ("SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_FMA_F64",
                 "SQ_INSTS_VALU_TRANS_F64")


def counters(db: str, names, trace_db: str | None = None, match: str = "") -> str:
    """Per-kernel averages per dispatch of the counters `names` (one
    --pmc run).  With the FP64 instruction counters and a kernel trace of
    the same command, also the FP64 operations per dispatch (64 lanes per
    wave instruction, an FMA counted as two) and their rate over the traced
    average duration: an upper bound where lanes are masked."""
    per = {n: _per_kernel(db, n) for n in names}
    dur = _durations(trace_db) if trace_db else {}
    kernels = sorted({k for d in per.values() for k in d if match in k},
                     key=lambda k: -max(per[n].get(k, (0, 0))[0] for n in names))
    f64 = all(n in names for n in FP64_COUNTERS)
    lds = "SQ_LDS_BANK_CONFLICT" in names and "SQ_INSTS_LDS" in names
    wait = "SQ_WAIT_INST_ANY" in names and "SQ_WAVE_CYCLES" in names
    head = f"{'kernel':64s} " + " ".join(f"{n.replace('SQ_', ''):>18s}" for n in names)
    if f64:
        head += f" {'fp64_ops':>14s} {'avg_us':>9s} {'GFLOP/s':>9s}"
    if lds:
        head += f" {'confl/lds':>9s}"
    if wait:
        head += f" {'wait_frac':>9s}"
    out = [head]
    for k in kernels:
        vals = [per[n].get(k, (0.0, 0))[0] for n in names]
        row = f"{k.replace(' ', '')[:64]:64s} " + " ".join(f"{v:18.0f}" for v in vals)
        if f64:
            v = dict(zip(names, vals))
            ops = 64.0 * (v["SQ_INSTS_VALU_ADD_F64"] + v["SQ_INSTS_VALU_MUL_F64"] +
                          v["SQ_INSTS_VALU_TRANS_F64"] + 2.0 * v["SQ_INSTS_VALU_FMA_F64"])
            d = dur.get(k, (0.0, 0))[0]
            row += f" {ops:14.4g} {d / 1e3:9.2f} " + (f"{ops / d:9.1f}" if d else f"{'-':>9s}")
        if lds:
            v = dict(zip(names, vals))
            n_lds = v["SQ_INSTS_LDS"]
            row += f" {v['SQ_LDS_BANK_CONFLICT'] / n_lds:9.2f}" if n_lds else f" {'-':>9s}"
        if wait:
            v = dict(zip(names, vals))
            cyc = v["SQ_WAVE_CYCLES"]
            row += f" {v['SQ_WAIT_INST_ANY'] / cyc:9.3f}" if cyc else f" {'-':>9s}"
        out.append(row)
    return "\n".join(out)


def read_traffic(path: str, kernel: str):
    """HBM bytes per dispatch (fetch + write) of `kernel` from a summary
    written by `traffic`, or None."""
    try:
        with open(path) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) == 7 and parts[0] == kernel:
                    return float(parts[5])
    except OSError:
        return None
    return None


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
    m.add_argument("--calib-fetch", help="FETCH_SIZE results of the calibration copy")
    m.add_argument("--calib-write", help="WRITE_SIZE results of the calibration copy")
    m.add_argument("--calib-kernel", default="vk::k_copy_interior")
    m.add_argument("--calib-bytes", type=float,
                   help="bytes the calibration copy reads (and writes) per launch")
    c = sub.add_parser("counters")
    c.add_argument("db")
    c.add_argument("names", nargs="+")
    c.add_argument("--trace-db", help="kernel trace of the same command, for durations")
    c.add_argument("--match", default="", help="only kernels whose name contains this")
    a = ap.parse_args()
    if a.cmd == "trace":
        print(trace(a.db, a.steps))
    elif a.cmd == "counters":
        print(counters(a.db, a.names, a.trace_db, a.match))
    else:
        fs = ws = 1.0
        if a.calib_fetch and a.calib_write and a.calib_bytes:
            fs, ws = calib_scale(a.calib_fetch, a.calib_write, a.calib_kernel, a.calib_bytes)
        print(traffic(a.fetch_db, a.write_db, a.points, fs, ws))


if __name__ == "__main__":
    main()
