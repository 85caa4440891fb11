"""The oracle's restatements of the reference's run-time diagnostics
(CHECK_COURANT, MONITOR, CHECK_CONVERGENCE, nse_dominik.f:2519-2892),
pinned where the reference can pin them:

- MONITOR returns everything it computes through its arguments, so the
  oracle is compared bit for bit with the reference's own compiled routine
  over two calls (the l = 1 reset and the differences after it);
- CHECK_CONVERGENCE appends l*dt and its sum to timely_convergence.dat;
  the value read back from that file pins the oracle's sum, and on the
  Poiseuille profile the sum has the closed form nz ny Re h^2 / 8;
- CHECK_COURANT only prints its result, so the oracle is pinned by the
  definition (a maximum of exactly rounded ratios, from 0) computed in
  numpy, with negative and mixed-sign fields that a sign slip or a
  dropped zero start would fail.
The position of MONITOR's largest |T - T_before| (printed by the
reference, returned by the oracle) is pinned by numpy's first-occurrence
argmax in the reference's k, j, i loop order.
"""
import ctypes
import math
import os

import numpy as np
import pytest

from oracle_bind import D, I, P, oracle, ref
from vtkmakecompare_amd import grid as G

needs_ref = pytest.mark.skipif(ref() is None, reason="reference library not built")


def _fields(g, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-1.0, 1.0, size=g.shape) for _ in range(5)]


def courant_numpy(g, dt, u, v, w):
    cx = u[1:-1, 1:-1, 1:-1] * dt / g.dx[1:-1][None, None, :]
    cy = v[1:-1, 1:-1, 1:-1] * dt / g.dy[1:-1][None, :, None]
    cz = w[1:-1, 1:-1, 1:-1] * dt / g.dz[1:-1][:, None, None]
    return max(0.0, float(cx.max()), float(cy.max()), float(cz.max()))


def o_courant(g, dt, u, v, w):
    r = ctypes.c_double(-1.0)
    oracle().o_check_courant(dt, P(g.dx), P(g.dy), P(g.dz), g.nx, g.ny, g.nz, P(u), P(v),
                             P(w), ctypes.byref(r))
    return r.value


@pytest.mark.parametrize("n", [(8, 9, 8), (12, 7, 6)])
def test_courant_is_the_signed_maximum(n):
    g = G.channel_grid(*n)
    u, v, w, _, _ = _fields(g, 11)
    dt = 1e-3
    assert o_courant(g, dt, u, v, w) == courant_numpy(g, dt, u, v, w)
    # every ratio negative: the maxima stay at their start, 0
    assert o_courant(g, dt, -np.abs(u) - 1.0, -np.abs(v) - 1.0, -np.abs(w) - 1.0) == 0.0
    # the largest ratio in w alone, at one point
    w2 = np.zeros(g.shape)
    w2[3, 2, 4] = 5.0
    assert o_courant(g, dt, np.zeros(g.shape), np.zeros(g.shape), w2) == 5.0 * dt / g.dz[3]


def o_monitor(g, l, t, tb, u, v, w, state):
    out = {k: ctypes.c_double(0.0) for k in ("dif", "t_all", "t_dif", "vel_dif", "u_dif",
                                            "v_dif", "w_dif")}
    loc = (ctypes.c_int * 3)()
    oracle().o_monitor(g.nx, g.ny, g.nz, l, P(t), P(g.dx), P(g.dy), P(g.dz),
                       ctypes.byref(out["dif"]), ctypes.byref(out["t_all"]),
                       ctypes.byref(out["t_dif"]), P(tb), P(u), P(v), P(w),
                       ctypes.byref(out["vel_dif"]), ctypes.byref(out["u_dif"]),
                       ctypes.byref(out["v_dif"]), ctypes.byref(out["w_dif"]),
                       ctypes.byref(state["u_int_before"]), ctypes.byref(state["v_int_before"]),
                       ctypes.byref(state["w_int_before"]), ctypes.byref(state["t_b"]), 1e-3,
                       loc)
    return {k: v.value for k, v in out.items()}, tuple(loc)


def r_monitor(g, l, t, tb, u, v, w, state):
    out = {k: ctypes.c_double(0.0) for k in ("dif", "t_all", "t_dif", "vel_dif", "u_dif",
                                            "v_dif", "w_dif")}
    ref().monitor_(I(g.nx), I(g.ny), I(g.nz), I(l), P(t), P(g.dx), P(g.dy), P(g.dz),
                   ctypes.byref(out["dif"]), ctypes.byref(out["t_all"]),
                   ctypes.byref(out["t_dif"]), P(tb), P(u), P(v), P(w),
                   ctypes.byref(out["vel_dif"]), ctypes.byref(out["u_dif"]),
                   ctypes.byref(out["v_dif"]), ctypes.byref(out["w_dif"]),
                   ctypes.byref(state["u_int_before"]), ctypes.byref(state["v_int_before"]),
                   ctypes.byref(state["w_int_before"]), ctypes.byref(state["t_b"]), D(1e-3),
                   I(0))
    return {k: v.value for k, v in out.items()}


def _state():
    return {k: ctypes.c_double(0.0) for k in ("u_int_before", "v_int_before", "w_int_before",
                                              "t_b")}


@needs_ref
@pytest.mark.parametrize("n", [(8, 9, 8), (12, 7, 6)])
def test_monitor_matches_reference(n):
    g = G.channel_grid(*n)
    u, v, w, t, tb = _fields(g, 5)
    so, sr = _state(), _state()
    for l, scale in ((1, 1.0), (2, 1.1)):
        o, _ = o_monitor(g, l, t * scale, tb, u * scale, v, w * scale, so)
        r = r_monitor(g, l, t * scale, tb, u * scale, v, w * scale, sr)
        assert o == r, (l, o, r)
        assert {k: x.value for k, x in so.items()} == {k: x.value for k, x in sr.items()}


def test_monitor_closed_forms():
    """T_all integrates the top halo plane k = nz+1 (the loop index after
    the k loop); the position is the first largest |T - T_before|."""
    g = G.channel_grid(10, 7, 6)
    u, v, w, t, tb = _fields(g, 9)
    t[1:-1, 1:-1, 1:-1] = 0.0
    tb[...] = 0.0
    t[2, 3, 4] = 2.0
    t[5, 1, 7] = -2.0  # the same |difference| later in k, j, i order
    t[-1] = 3.0        # the plane k = nz + 1
    o, loc = o_monitor(g, 1, t, tb, u, v, w, _state())
    assert loc == (4, 3, 2)
    assert o["t_dif"] == 2.0
    expect = math.fsum(3.0 * g.dx[i] * g.dy[j] * g.dz[g.nz + 1]
                       for j in range(1, g.ny + 1) for i in range(1, g.nx + 1))
    assert o["t_all"] == pytest.approx(expect, rel=1e-14)
    assert o["dif"] == 1.0  # |(T_all - 0) / T_all| at l = 1
    d = np.abs(np.random.default_rng(1).uniform(-1, 1, size=(g.nz, g.ny, g.nx)))
    t2 = np.zeros(g.shape)
    t2[1:-1, 1:-1, 1:-1] = d
    _, loc2 = o_monitor(g, 1, t2, np.zeros(g.shape), u, v, w, _state())
    kk, jj, ii = np.unravel_index(np.argmax(d), d.shape)
    assert loc2 == (ii + 1, jj + 1, kk + 1)


def o_convergence(g, u, re):
    r = ctypes.c_double(-1.0)
    oracle().o_check_convergence(P(g.y), g.yl, P(u), re, g.nx, g.ny, g.nz, ctypes.byref(r))
    return r.value


@needs_ref
def test_convergence_matches_reference_file(tmp_path):
    g = G.channel_grid(8, 9, 6)
    u = np.random.default_rng(2).uniform(-3.0, 3.0, size=g.shape)
    cwd = os.getcwd()
    os.chdir(tmp_path)  # the reference appends to timely_convergence.dat
    try:
        ref().check_convergence_(P(g.y), D(g.yl), P(u), I(7), D(1e-3), D(20.0), I(g.nx),
                                 I(g.ny), I(g.nz))
        rec = open("timely_convergence.dat").read().split()
    finally:
        os.chdir(cwd)
    assert float(rec[0]) == pytest.approx(7e-3, rel=1e-15)
    assert o_convergence(g, u, 20.0) == pytest.approx(float(rec[1]), rel=1e-15)


def test_convergence_poiseuille_closed_form():
    """On the uniform grid the discrete profile u* = -Re/2 y (y - yl) +
    Re h^2/8 is off the parabola by Re h^2/8 at every point."""
    re = 100.0
    g = G.set_c(8, 16, 6, 2.0 * math.pi, 2.0, math.pi)
    h = g.yl / g.ny
    yc = g.y[1: g.ny + 1]
    u = np.zeros(g.shape)
    u[1:-1, 1:-1, 1:-1] = (-re / 2.0 * yc * (yc - g.yl) + re * h * h / 8.0)[None, :, None]
    assert o_convergence(g, u, re) == pytest.approx(g.nz * g.ny * re * h * h / 8.0, rel=1e-12)
