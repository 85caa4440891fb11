"""Benchmark: the fused time step of the channel solver on one MI355X, in
its FP64 form and its mixed-precision form.

Workload: step384 -- 384 x 193 x 384 channel (Lx = 2 pi, Lz = pi, walls at
y = +-1 with tanh stretching beta = 1.8), Re = 5600, Pr = 0.71, 8 seeded
spheres as the level set (set_levelset on the device), seeded +-0.05
noise; synthetic inputs, every field resident in HBM before timing.
One step is one pass of the reference's time-loop body
(nse_dominik.f:290-448): halos, explicit terms with penalisation and AB2,
the nine ADI line sweeps plus three for temperature, the spectral
pressure projection (with one FP64 refinement pass in the mixed form).

Both forms start from the same state and run W + K steps; the line
reports the mixed form's throughput as `value` (M grid-point-steps/s),
the FP64 form's beside it, the ratio, the mixed form's normwise error
against the FP64 form per field after the same steps, the post-projection
divergence of each, the roofline of each form's dominant kernel
(algorithmic bytes per launch over its average duration, timed with HIP
events on the library stream during the timed steps) and the oracle's CPU
time on a bounded sample.
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

# device symbol of each timed kernel, as rocprofv3 names it (the line
# sweeps are two kernels each -- the recorded-multiplier kernel and the
# general kernel for blocks next to the solid -- and are reported by their
# sweep name only)
KERNEL_SYMBOL = {
    "explicit_vel": "vk::k_explicit_fused<1>",
    "explicit_t": "vk::k_explicit_fused<2>",
    "explicit_mixed": "vk::k_explicit_mixed<1>",
    "explicit_mixed_t": "vk::k_explicit_mixed<2>",
    "project": "vk::k_project<false>",
    "project_mixed": "vk::k_project_mixed",
    "rfft_x_div": "vk::k_rfft_x<vk::SampleDiv,double>",
    "irfft_x": "vk::k_irfft_x<double>",
}


def kernel_bytes(name: str, g, with_ep2: bool = False, mixed: bool = False) -> float:
    """Algorithmic HBM bytes of one launch of a kernel of the fused step:
    each array the kernel must read or write counted once over the
    interior (halos ignored).  FP64 form: 8 bytes per double, 16 per
    complex.  Mixed form: the velocity, u_temp, p and divu in double, the
    temperature, history, u_ta, cellvalue, rhoCp, thc, the spectra and
    the Poisson solutions in float."""
    N = g.nx * g.ny * g.nz
    f = 8.0 * N
    if not mixed:
        S = (g.nx // 2 + 1) * g.ny * g.nz * 16.0
        table = {
            # u,v,w,T,rhoCp,thc, 4 history, 3 u_ta, 4 cellvalue in; 4 rhs, 4 history out
            "explicit_fused": 25 * f,
            # the same in two kernels: u,v,w, 3 history, 3 u_ta, cellvalue 2-4 in,
            # 3 rhs and 3 history out; u,v,w, T, rhoCp, thc, history, cellvalue 1
            # in, T work and history out
            "explicit_vel": 18 * f, "explicit_t": 10 * f,
            "line_y_face": 3 * f, "line_y_v": 3 * f, "line_z_face": 3 * f, "line_x_face": 3 * f,
            # T lines also read rhoCp and thc
            "line_y_t": 5 * f, "line_x_t": 5 * f, "line_z_t": 5 * f,
            "rfft_x_div": 3 * f + S, "rfft_x": f + S,
            "cfft_z_fwd": 2 * S, "cfft_z_inv": 2 * S, "line_spectral": 2 * S,
            "irfft_x": S + f,
            "project": (18 if with_ep2 else 12) * f,
        }
        return table.get(name, 0.0)
    h = 4.0 * N  # one float field
    S = (g.nx // 2 + 1) * g.ny * g.nz * 8.0
    table = {
        # velocity part -- in: u,v,w (double), 3 history, 3 u_ta, cellvalue 2-4
        # (float); out: u_temp, v_temp, w_temp (double), 3 history (float).
        # temperature part -- in: u,v,w (double), T, rhoCp, thc, history,
        # cellvalue 1 (float); out: T work, history (float)
        "explicit_mixed": 6 * f + 12 * h, "explicit_mixed_t": 3 * f + 7 * h,
        # velocity lines: double rhs and solution, float cellvalue
        "line_y_face": 2 * f + h, "line_y_v": 2 * f + h, "line_z_face": 2 * f + h,
        "line_x_face": 2 * f + h,
        # temperature lines: all float (rhs, cellvalue, rhoCp, thc, solution)
        "line_y_t": 5 * h, "line_x_t": 5 * h, "line_z_t": 5 * h,
        # pass 1 keeps the FP64 right-hand side for the residual of pass 2
        "rfft_x_div": 3 * f + f + S, "rfft_x_resid": f + h + S,
        "cfft_z_fwd": 2 * S, "cfft_z_fwd2": 2 * S, "cfft_z_inv": 2 * S, "cfft_z_inv2": 2 * S,
        "line_spectral": 2 * S,
        "irfft_x": S + h, "irfft_x2": S + h,
        # phi1, delta, u_temp x3 in; u, v, w, p, divu (double), 3 u_ta (float) out
        "project_mixed": 2 * h + 3 * f + 5 * f + 3 * h,
    }
    return table.get(name, 0.0)


STEP_LAUNCHES = {"explicit_vel": 1, "explicit_t": 1, "explicit_mixed": 1, "explicit_mixed_t": 1, "line_y_face": 2, "line_y_v": 1,
                 "line_z_face": 3, "line_x_face": 3, "line_y_t": 1, "line_x_t": 1,
                 "line_z_t": 1, "rfft_x_div": 1, "rfft_x_resid": 1, "cfft_z_fwd": 1,
                 "cfft_z_fwd2": 1, "cfft_z_inv": 1, "cfft_z_inv2": 1, "irfft_x": 1,
                 "irfft_x2": 1, "project": 1, "project_mixed": 1}


def step_bytes(g, with_ep2: bool = False, mixed: bool = False) -> float:
    per_step = dict(STEP_LAUNCHES)
    per_step["line_spectral"] = 2 if mixed else 1
    return sum(c * kernel_bytes(k, g, with_ep2, mixed) for k, c in per_step.items())


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


PROFILED_STEPS = 10


def timed_run(lib, run, steps: int, warmup: int):
    """W untimed steps, then K timed steps (the step's kernels spread over
    the library's streams as in production), then PROFILED_STEPS more with
    per-kernel HIP-event timing, during which the library issues every
    kernel on one stream so that each kernel is timed alone."""
    from vtkmakecompare_amd import _lib, api
    run.step(warmup)
    api.synchronize()
    t0 = time.perf_counter()
    run.step(steps)
    api.synchronize()
    t1 = time.perf_counter()
    # --steps 0 (for kernel traces): the profiled steps only, no throughput
    el = (t1 - t0) if steps > 0 else None
    _lib.check(lib.vk_profile_begin(), "profile")
    run.step(PROFILED_STEPS)
    api.synchronize()
    _lib.check(lib.vk_profile_end(), "profile")
    return el, profile_entries(lib)


def diagnostics_timing(lib, run, cfg, g, reps: int = 10):
    """CHECK_COURANT, MONITOR and CHECK_CONVERGENCE on the FP64 run's
    resident fields (T_pre standing in for T_before), each timed with HIP
    events around its two passes.  Algorithmic bytes: u, v, w for the
    Courant maximum; u, v, w, T, T_before and the plane k = nz+1 of T for
    MONITOR; the plane i = nx of u for the convergence sum."""
    from vtkmakecompare_amd import _lib, api
    d = run.d
    st = api.MonitorState()
    # one untimed round first: the first launch of a module's kernels loads
    # its code object
    for r in range(reps + 1):
        if r == 1:
            api.synchronize()
            _lib.check(lib.vk_profile_begin(), "profile")
        api.check_courant(cfg.dt, d["dx"], d["dy"], d["dz"], g.nx, g.ny, g.nz, d["u"], d["v"],
                          d["w"])
        api.monitor(g.nx, g.ny, g.nz, r + 1, d["t"], d["dx"], d["dy"], d["dz"], d["t_pre"],
                    d["u"], d["v"], d["w"], cfg.dt, st)
        api.check_convergence(d["y"], g.yl, d["u"], r + 1, cfg.dt, cfg.re, g.nx, g.ny, g.nz)
    api.synchronize()
    _lib.check(lib.vk_profile_end(), "profile")
    prof = profile_entries(lib)
    n = g.npoints
    alg = {"check_courant": 3 * 8 * n, "monitor": 5 * 8 * n + 8 * g.nx * g.ny,
           "check_convergence": 8 * g.ny * g.nz}
    out = {}
    for k, b in alg.items():
        if k in prof:
            ms, cnt = prof[k]
            avg_s = ms / cnt / 1e3
            out[k] = {"avg_us": round(avg_s * 1e6, 2), "launches": cnt, "alg_bytes": b,
                      "alg_GBs": round(b / avg_s / 1e9, 1)}
    return out


def kernel_table(prof, g, with_ep2, mixed):
    return {k: {"total_ms": round(v[0], 3), "launches": v[1],
                "avg_us": round(1e3 * v[0] / v[1], 2),
                "alg_GBs": round(kernel_bytes(k, g, with_ep2, mixed) / (v[0] / v[1] / 1e3) / 1e9, 1)
                if kernel_bytes(k, g, with_ep2, mixed) else None}
            for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0])}


def roofline(prof, g, with_ep2, mixed, traffic_file):
    """The dominant kernel (by total time): algorithmic bytes per launch
    over its average launch duration, against the 8 TB/s HBM peak."""
    from vtkmakecompare_amd.profsummary import read_traffic
    dname, (dms, dcount) = max(prof.items(), key=lambda kv: kv[1][0])
    avg_s = dms / dcount / 1e3
    alg = kernel_bytes(dname, g, with_ep2, mixed)
    traffic = None
    if dname in KERNEL_SYMBOL and not with_ep2:
        traffic = read_traffic(traffic_file, KERNEL_SYMBOL[dname])
    achieved = alg / avg_s / 1e9 if alg else 0.0
    return {"bound": "hbm", "kernel": dname, "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "alg_bytes_per_launch": alg, "avg_launch_us": round(avg_s * 1e6, 2),
            "traffic": traffic}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200,
                    help="timed steps (0: only the profiled steps, for kernel traces)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="step384")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=("both", "fp64", "mixed"), default="both",
                    help="run one form only (for profiling)")
    ap.add_argument("--diagnostics", action="store_true",
                    help="also compute MKP's u_ep2 splitting-error fields (FP64 form)")
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
    npts = g.npoints
    print(f"[bench] {cfg.name} {g.nx}x{g.ny}x{g.nz}: level set on device", file=sys.stderr,
          flush=True)
    # the level-set setup (set_levelset's 2001 reinitialisation iterations)
    # timed on its own: HIP events around the iteration loop
    _lib.check(lib.vk_profile_begin(), "profile")
    r64 = api.ChannelRun(cfg, g, st, diagnostics=args.diagnostics)
    api.synchronize()
    _lib.check(lib.vk_profile_end(), "profile")
    lsp = profile_entries(lib).get("levelset_reinit")
    levelset = None
    if lsp:
        per_it = lsp[0] / 2001 / 1e3
        # per iteration: phi (read with its neighbours) and phi0 in, phi out
        lb = 3 * 8 * npts
        levelset = {"total_ms": round(lsp[0], 2), "iterations": 2001,
                    "us_per_iteration": round(per_it * 1e6, 2), "alg_bytes_per_iteration": lb,
                    "alg_GBs": round(lb / per_it / 1e9, 1)}
    cv = r64.get("cellvalue")
    res = {}
    diag = None
    if args.only in ("both", "fp64"):
        el, prof = timed_run(lib, r64, args.steps, args.warmup)
        res["fp64"] = (el, prof, r64)
        diag = diagnostics_timing(lib, r64, cfg, g)
        if el:
            print(f"[bench] fp64: {1e3 * el / args.steps:.3f} ms/step", file=sys.stderr, flush=True)
    if args.only in ("both", "mixed"):
        if args.only == "mixed":
            r64 = None
            res.pop("fp64", None)
        rmx = api.ChannelRun(cfg, g, st, cellvalue=cv, precision="mixed")
        el, prof = timed_run(lib, rmx, args.steps, args.warmup)
        res["mixed"] = (el, prof, rmx)
        if el:
            print(f"[bench] mixed: {1e3 * el / args.steps:.3f} ms/step", file=sys.stderr,
                  flush=True)

    forms = {}
    for name, (el, prof, run) in res.items():
        mixed = name == "mixed"
        with_ep2 = args.diagnostics and not mixed
        ms = 1e3 * el / args.steps if el else None
        sb = step_bytes(g, with_ep2, mixed)
        tfile = os.path.join(ROOT, "profiles", f"{cfg.name}_traffic.txt")
        forms[name] = {
            "value": round(npts * args.steps / el / 1e6, 2) if el else None,
            "ms_per_step": round(ms, 4) if ms else None,
            "roofline": roofline(prof, g, with_ep2, mixed, tfile),
            "step_roofline": {"alg_bytes_per_step": sb,
                              "alg_GBs": round(sb / (ms / 1e3) / 1e9, 1) if ms else None,
                              "frac": round(sb / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
                              if ms else None,
                              # the kernels one after another (profiled steps);
                              # ms_per_step is less where streams overlap them
                              "kernel_time_ms_per_step":
                                  round(sum(v[0] for v in prof.values()) / PROFILED_STEPS, 4)},
            "kernels": kernel_table(prof, g, with_ep2, mixed),
            "max_abs_divu": api.field_absmax(run.d["divu"], g.nx, g.ny, g.nz),
        }

    # error of the mixed form against the FP64 form after the same steps
    errors = None
    if "fp64" in res and "mixed" in res:
        a, b = res["mixed"][2], res["fp64"][2]
        errors = {}
        for k in ("u", "v", "w", "p", "t"):
            x = a.get(k)[1:-1, 1:-1, 1:-1]
            y = b.get(k)[1:-1, 1:-1, 1:-1]
            errors[k] = float(np.max(np.abs(x - y)) / max(float(np.max(np.abs(y))), 1e-300))
            del x, y
        errors["velocity"] = max(errors["u"], errors["v"], errors["w"])
        errors["steps_compared"] = args.warmup + args.steps + PROFILED_STEPS

    cpu = None
    if not args.no_cpu_baseline:
        print(f"[bench] CPU baseline {args.cpu_steps} steps", file=sys.stderr, flush=True)
        st2 = G.initial_state(cfg, g)
        cpu = cpu_baseline(cfg, g, st2, cv, args.cpu_steps)

    head = forms.get("mixed") or forms["fp64"]
    is_mixed = "mixed" in forms
    out = {
        "metric": "M grid-pt-steps/s @384x193x384: mixed vs FP64 at <=1e-6 rel err; HBM GB/s",
        "value": head["value"],
        "unit": "M grid-pt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 (mixed)" if is_mixed else "f64",
        "data": "synthetic (seeded counter-based noise, 8 random spheres)",
        "config": {"workload": f"{cfg.name}: full time step, {g.nx}x{g.ny}x{g.nz} channel, "
                               f"Re={cfg.re:g}, 8 spheres, tanh-stretched y; value is the "
                               f"{'mixed-precision' if is_mixed else 'FP64'} step",
                   "precision": "mixed" if is_mixed else "fp64", "grid": [g.nx, g.ny, g.nz],
                   "mkp_diagnostics": args.diagnostics},
        "roofline": head["roofline"],
        "mixed_vs_fp64": None,
        "errors_mixed_vs_fp64": errors,
        "forms": forms,
        "diagnostics": diag,
        "levelset": levelset,
        "cpu_baseline": cpu,
    }
    if "fp64" in forms and "mixed" in forms and forms["fp64"]["value"]:
        out["mixed_vs_fp64"] = round(forms["mixed"]["value"] / forms["fp64"]["value"], 4)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
