"""Benchmark: the fused time step of the channel solver on one MI355X.

Workload: step384 -- 384 x 193 x 384 channel (Lx = 2 pi, Lz = pi, walls at
y = +-1 with tanh stretching beta = 1.8), Re = 5600, Pr = 0.71, 8 seeded
spheres as the level set (set_levelset on the device), seeded +-0.05
noise; synthetic inputs, every field resident in HBM before timing.
One step is one pass of the reference's time-loop body
(nse_dominik.f:290-448): halos, explicit terms with penalisation and AB2,
the nine ADI line sweeps plus three for temperature, the spectral
pressure projection.

Prints one JSON line with the throughput in M grid-point-steps/s, the
roofline of the dominant kernel (algorithmic bytes per launch over its
average duration, timed with HIP events on the library stream during the
timed steps) and the oracle's CPU time on a bounded sample.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# device symbol of each timed kernel of the fused step, as rocprofv3 names it
KERNEL_SYMBOL = {
    "explicit_fused": "vk::k_explicit_fused",
    "project": "vk::k_project",
    "line_y_face": "vk::k_line_solve<32,8,512,true,vk::SysYFace>",
    "line_y_v": "vk::k_line_solve<32,8,512,true,vk::SysYV>",
    "line_y_t": "vk::k_line_solve<32,8,512,true,vk::SysYT>",
    "line_z_t": "vk::k_line_solve<64,8,512,true,vk::SysPeriodicT>",
    "line_x_t": "vk::k_line_solve<64,8,256,false,vk::SysPeriodicT>",
    "line_spectral": "vk::k_line_solve<32,8,512,true,vk::SysSpectral>",
    "rfft_x_div": "vk::k_rfft_x<vk::SampleDiv>",
    "irfft_x": "vk::k_irfft_x",
}


def kernel_bytes(name: str, g, with_ep2: bool) -> float:
    """Algorithmic HBM bytes of one launch of a kernel of the fused step:
    each array the kernel must read or write counted once over the
    interior (halos ignored), 8 bytes per double, 16 per complex."""
    N = g.nx * g.ny * g.nz
    S = (g.nx // 2 + 1) * g.ny * g.nz * 16.0
    f = 8.0 * N
    table = {
        # u,v,w,T,rhoCp,thc, 4 history, 3 u_ta, 4 cellvalue in; 4 rhs, 4 history out
        "explicit_fused": 25 * f,
        "line_y_face": 3 * f, "line_y_v": 3 * f, "line_z_face": 3 * f, "line_x_face": 3 * f,
        # T lines also read rhoCp and thc
        "line_y_t": 5 * f, "line_x_t": 5 * f, "line_z_t": 5 * f,
        "rfft_x_div": 3 * f + S, "rfft_x": f + S,
        "cfft_z_fwd": 2 * S, "cfft_z_inv": 2 * S, "line_spectral": 2 * S,
        "irfft_x": S + f,
        "project": (18 if with_ep2 else 12) * f,
        "copy_interior": 2 * f,
    }
    return table.get(name, 0.0)


def step_bytes(g, with_ep2: bool) -> float:
    per_step = {"explicit_fused": 1, "line_y_face": 2, "line_y_v": 1, "line_z_face": 3,
                "line_x_face": 3, "line_y_t": 1, "line_x_t": 1, "line_z_t": 1,
                "rfft_x_div": 1, "cfft_z_fwd": 1, "cfft_z_inv": 1, "line_spectral": 1,
                "irfft_x": 1, "project": 1, "copy_interior": 1}
    return sum(c * kernel_bytes(k, g, with_ep2) for k, c in per_step.items())


def profile_entries(lib):
    n = ctypes.c_int(0)
    lib.vk_profile_count(ctypes.byref(n))
    out = {}
    name = ctypes.create_string_buffer(128)
    for q in range(n.value):
        ms = ctypes.c_double(0.0)
        cnt = ctypes.c_int(0)
        rc = lib.vk_profile_entry(ctypes.byref(ctypes.c_int(q)), name,
                                  ctypes.byref(ctypes.c_int(128)), ctypes.byref(ms),
                                  ctypes.byref(cnt))
        if rc == 0 and cnt.value > 0:
            out[name.value.decode()] = (ms.value, cnt.value)
    return out


def cpu_baseline(cfg, g, st, cellvalue, steps: int):
    """The oracle (C restatement, OpenMP) on the same workload for `steps`
    steps from the same initial state."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import OracleRun
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    st_o = dict(st)
    st_o["cellvalue"] = cellvalue
    run = OracleRun(cfg, g, st_o)
    t0 = time.perf_counter()
    run.step(steps)
    dt = time.perf_counter() - t0
    return {"value": g.npoints * steps / dt / 1e6, "unit": "M grid-pt-steps/s",
            "cores": threads, "kind": "port",
            "sample": f"{steps} full time steps of {cfg.name} "
                      f"({g.nx}x{g.ny}x{g.nz}), oracle C/OpenMP, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="step384")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--diagnostics", action="store_true",
                    help="also compute MKP's u_ep2 splitting-error fields")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU y-slab design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from vtkmakecompare_amd import _lib, api
    from vtkmakecompare_amd import grid as G

    api.init(0)
    lib = _lib.lib()
    cfg = G.make_config(args.config)
    g = cfg.grid()
    st = G.initial_state(cfg, g)
    print(f"[bench] {cfg.name} {g.nx}x{g.ny}x{g.nz}: level set on device", file=sys.stderr,
          flush=True)
    run = api.ChannelRun(cfg, g, st, diagnostics=args.diagnostics)
    api.synchronize()
    run.step(args.warmup)
    api.synchronize()
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)

    _lib.check(lib.vk_profile_begin(), "profile")
    t0 = time.perf_counter()
    run.step(args.steps)
    api.synchronize()
    t1 = time.perf_counter()
    _lib.check(lib.vk_profile_end(), "profile")
    prof = profile_entries(lib)
    elapsed = t1 - t0
    npts = g.npoints
    value = npts * args.steps / elapsed / 1e6
    ms_step = 1e3 * elapsed / args.steps

    # divergence check of the final state (post-projection, absolute)
    div = api.field_absmax(run.d["divu"], g.nx, g.ny, g.nz)
    umax = api.field_absmax(run.d["u"], g.nx, g.ny, g.nz)

    # dominant kernel by total time
    with_ep2 = args.diagnostics
    dom = max(prof.items(), key=lambda kv: kv[1][0])
    dname, (dms, dcount) = dom
    avg_s = dms / dcount / 1e3
    alg = kernel_bytes(dname, g, with_ep2)
    # HBM bytes per launch from the committed FETCH_SIZE + WRITE_SIZE summary
    # of this workload (two rocprofv3 --pmc runs, scaled by a known copy)
    from vtkmakecompare_amd.profsummary import read_traffic
    traffic = None
    if cfg.name == "step384" and not with_ep2 and dname in KERNEL_SYMBOL:
        traffic = read_traffic(os.path.join(ROOT, "profiles", "step384_traffic.txt"),
                               KERNEL_SYMBOL[dname])
    achieved = alg / avg_s / 1e9 if alg else 0.0
    kernels = {k: {"total_ms": round(v[0], 3), "launches": v[1],
                   "avg_us": round(1e3 * v[0] / v[1], 2),
                   "alg_GBs": round(kernel_bytes(k, g, with_ep2) / (v[0] / v[1] / 1e3) / 1e9, 1)
                   if kernel_bytes(k, g, with_ep2) else None}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0])}
    kernel_sum_ms = sum(v[0] for v in prof.values()) / args.steps

    cpu = None
    if not args.no_cpu_baseline:
        print(f"[bench] GPU done ({ms_step:.3f} ms/step); CPU baseline "
              f"{args.cpu_steps} steps", file=sys.stderr, flush=True)
        cv = run.get("cellvalue")
        st2 = G.initial_state(cfg, g)
        cpu = cpu_baseline(cfg, g, st2, cv, args.cpu_steps)

    out = {
        "metric": "M grid-pt-steps/s @384x193x384: mixed vs FP64 at <=1e-6 rel err; HBM GB/s",
        "value": round(value, 2),
        "unit": "M grid-pt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded counter-based noise, 8 random spheres)",
        "config": {"workload": f"{cfg.name}: full FP64 time step, {g.nx}x{g.ny}x{g.nz} "
                               f"channel, Re={cfg.re:g}, 8 spheres, tanh-stretched y",
                   "precision": "fp64", "grid": [g.nx, g.ny, g.nz],
                   "mkp_diagnostics": with_ep2},
        "roofline": {"bound": "hbm", "kernel": dname, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "alg_bytes_per_launch": alg, "avg_launch_us": round(avg_s * 1e6, 2),
                     "traffic": traffic},
        "step_roofline": {"alg_bytes_per_step": step_bytes(g, with_ep2),
                          "alg_GBs": round(step_bytes(g, with_ep2) / (ms_step / 1e3) / 1e9, 1),
                          "frac": round(step_bytes(g, with_ep2) / (ms_step / 1e3) / 1e9
                                        / HBM_PEAK_GBS, 4),
                          "kernel_time_ms_per_step": round(kernel_sum_ms, 4)},
        "kernels": kernels,
        "check": {"max_abs_divu": div, "max_abs_u": umax},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
