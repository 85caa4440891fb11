"""The large mixed-precision forward and adjoint run, against the FP64
adjoint of the same problem.

    python -m vtkmakecompare_amd.large_adjoint [--config large768] [--steps 400]
        [--snapshots -1] [--directions 4] [--fd-directions 1] [--eps 1e-6]

Workload large768: 768 x 385 x 768 channel (227,082,240 points), Re 5600,
Pr 0.71, 64 seeded rotated ellipsoids as the level set (generated and
reinitialised on the device), seeded +-0.05 noise; 400 steps.  The
objective is the wall heat flux integrated over the walls and the run;
the gradient is dJ/dcellvalue(...,1).

Both gradients use binomial checkpointing (vk_adjoint_gradient_revolve)
with as many snapshots as the device memory holds:
- mixed: the forward trajectory of the mixed-precision step (T, the
  history, u_ta, cellvalue, rhoCp, thc in float; snapshots 3 double and 8
  float fields), the FP64 adjoint linearised about it;
- FP64: the FP64 step's trajectory and its exact discrete adjoint.
Reported: both objectives, the mixed gradient's error against the FP64
adjoint (normwise, in L2, and on seeded random directions), a central
finite difference of the FP64 objective on --fd-directions directions,
the time of each gradient, and the snapshots, forward steps and device
memory each took.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import api
from . import grid as G


def log(msg):
    print(f"[large_adjoint] {msg}", file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="large768")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--snapshots", type=int, default=-1)
    ap.add_argument("--directions", type=int, default=4)
    ap.add_argument("--fd-directions", type=int, default=1)
    ap.add_argument("--eps", type=float, default=1e-6)
    a = ap.parse_args(argv)
    os.environ.setdefault("VK_ADJOINT_PROGRESS", "1")
    api.init(0)
    cfg = G.make_config(a.config)
    g = cfg.grid()
    nsteps = a.steps or cfg.steps
    log(f"{cfg.name} {g.nx}x{g.ny}x{g.nz}, {nsteps} steps: initial state")
    st = G.initial_state(cfg, g, lean=True)
    log("level set on the device")
    phi0 = api.device_phi0(cfg, g, st)
    cv = api.levelset_cellvalue(g, phi0)
    phi0.free()
    prob = api.AdjointProblem(cfg, g, st, cv, nsteps)
    del st
    shapes = (f"{len(cfg.ellipsoids)} ellipsoids" if cfg.ellipsoids is not None
              else f"{len(cfg.spheres)} spheres")
    out = {"workload": f"{cfg.name}: {g.nx}x{g.ny}x{g.nz} ({g.npoints} points), Re "
                       f"{cfg.re:g}, {shapes}, {nsteps} steps, wall-heat-flux objective, "
                       f"gradient with respect to cellvalue(...,1)"}
    res = {}
    for form in ("mixed", "fp64"):
        log(f"{form} gradient")
        t0 = time.perf_counter()
        j, grad, stats = prob.gradient_revolve(snapshots=a.snapshots, mixed=form == "mixed")
        el = time.perf_counter() - t0
        res[form] = (j, grad)
        stats["device_GB"] = round(stats.pop("device_bytes") / 1e9, 1)
        out[form] = {"objective": j, "seconds": round(el, 1), **stats}
        log(f"{form}: J = {j:.15e} in {el:.1f} s, {stats}")
    (jm, gm), (jd, gd) = res["mixed"], res["fp64"]
    scale = float(np.max(np.abs(gd)))
    out["objective_rel_err"] = abs(jm - jd) / abs(jd)
    out["gradient_rel_err_max"] = float(np.max(np.abs(gm - gd))) / scale
    out["gradient_rel_err_l2"] = float(np.linalg.norm((gm - gd).ravel()) /
                                       np.linalg.norm(gd.ravel()))
    rng = np.random.default_rng(cfg.seed)
    dirs = []
    for q in range(a.directions + a.fd_directions):
        d = rng.uniform(-1.0, 1.0, size=gd.shape)
        row = {"mixed": float(np.sum(gm * d)), "fp64": float(np.sum(gd * d))}
        row["rel"] = abs(row["mixed"] - row["fp64"]) / abs(row["fp64"])
        if q >= a.directions:  # central difference of the FP64 objective
            log(f"finite difference on direction {q}")
            cp = cv.copy()
            cm = cv.copy()
            cp[0] += a.eps * d
            cm[0] -= a.eps * d
            fd = (prob.objective(cp) - prob.objective(cm)) / (2.0 * a.eps)
            row["fd_fp64"] = fd
            row["fd_rel"] = abs(row["fp64"] - fd) / abs(fd)
            del cp, cm
        dirs.append(row)
        log(f"direction {q}: {row}")
    out["directions"] = dirs
    out["eps"] = a.eps
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
