"""The run-time diagnostics on the device against the oracle:
CHECK_COURANT (nse_dominik.f:2519), MONITOR (:2797) and CHECK_CONVERGENCE
(:2875).  Maxima and MONITOR's reported position are exact; the sums are
reduced in a tree instead of the reference's serial loop, so they agree to
rounding (1e-13 relative here: the sums hold at most a few hundred
thousand positive terms, whose reordering moves them by a few ulp).
"""
import ctypes
import math

import numpy as np
import pytest

from oracle_bind import P, oracle
from vtkmakecompare_amd import grid as G

pytestmark = pytest.mark.gpu

SIZES = [(24, 17, 12), (64, 33, 48)]
SUM_TOL = 1e-13


@pytest.fixture(scope="module")
def api():
    from vtkmakecompare_amd import api as A
    A.init()
    return A


def dev(a):
    from vtkmakecompare_amd.device import DeviceArray
    return DeviceArray.from_host(np.ascontiguousarray(a))


def metrics(g):
    return {k: dev(getattr(g, k)) for k in ("dx", "dy", "dz", "y")}


def fields(g, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-1.0, 1.0, size=g.shape) for _ in range(5)]


def o_courant(g, dt, u, v, w):
    r = ctypes.c_double(-1.0)
    oracle().o_check_courant(dt, P(g.dx), P(g.dy), P(g.dz), g.nx, g.ny, g.nz, P(u), P(v),
                             P(w), ctypes.byref(r))
    return r.value


@pytest.mark.parametrize("n", SIZES)
def test_courant_exact(api, n):
    g = G.channel_grid(*n)
    m = metrics(g)
    u, v, w, _, _ = fields(g, 21)
    dt = 1e-3
    # quotients within a few ulp of 0.7 everywhere: the maximum must be the
    # exact largest quotient, so a reduction that loses or reorders a
    # near-tie shows here as a last-bit difference
    rng = np.random.default_rng(5)
    ulps = 1.0 + rng.integers(-6, 7, size=g.shape) * 2.0 ** -52
    un = 0.7 * (g.dx / dt)[None, None, :] * ulps
    vn = 0.7 * (g.dy / dt)[None, :, None] * ulps[:, ::-1, :]
    wn = 0.7 * (g.dz / dt)[:, None, None] * ulps[::-1, :, :]
    for a, b, c in ((u, v, w), (-np.abs(u), -np.abs(v), -np.abs(w)), (u, 3.0 * v, w),
                    (un, vn, wn)):
        got = api.check_courant(dt, m["dx"], m["dy"], m["dz"], g.nx, g.ny, g.nz, dev(a),
                                dev(b), dev(c))
        assert got == o_courant(g, dt, a, b, c)


def o_monitor(g, l, t, tb, u, v, w, st):
    out = {k: ctypes.c_double(0.0) for k in ("dif", "t_all", "t_dif", "vel_dif", "u_dif",
                                            "v_dif", "w_dif")}
    loc = (ctypes.c_int * 3)()
    oracle().o_monitor(g.nx, g.ny, g.nz, l, P(t), P(g.dx), P(g.dy), P(g.dz),
                       ctypes.byref(out["dif"]), ctypes.byref(out["t_all"]),
                       ctypes.byref(out["t_dif"]), P(tb), P(u), P(v), P(w),
                       ctypes.byref(out["vel_dif"]), ctypes.byref(out["u_dif"]),
                       ctypes.byref(out["v_dif"]), ctypes.byref(out["w_dif"]),
                       ctypes.byref(st["u_int_before"]), ctypes.byref(st["v_int_before"]),
                       ctypes.byref(st["w_int_before"]), ctypes.byref(st["t_b"]), 1e-3, loc)
    r = {k: x.value for k, x in out.items()}
    r.update({k: x.value for k, x in st.items()})
    return r, tuple(loc)


@pytest.mark.parametrize("n", SIZES)
def test_monitor_against_oracle(api, n):
    g = G.channel_grid(*n)
    m = metrics(g)
    u, v, w, t, tb = fields(g, 8)
    # two equal largest differences: the first in k, j, i order is reported
    t[3, 5, 7] = tb[3, 5, 7] + 4.0
    t[6, 2, 1] = tb[6, 2, 1] - 4.0
    ost = {k: ctypes.c_double(0.0) for k in ("u_int_before", "v_int_before", "w_int_before",
                                             "t_b")}
    gst = api.MonitorState()
    # the second call scales by 2 (exact), so the tie and its position stay
    for l, s in ((1, 1.0), (2, 2.0)):
        ta, tba, ua, wa = t * s, tb * s, u * s, w / s
        o, oloc = o_monitor(g, l, ta, tba, ua, v, wa, ost)
        api.monitor(g.nx, g.ny, g.nz, l, dev(ta), m["dx"], m["dy"], m["dz"], dev(tba),
                    dev(ua), dev(v), dev(wa), 1e-3, gst)
        assert gst.loc == oloc == (7, 5, 3)  # t[k=3, j=5, i=7] comes first
        assert gst["t_dif"] == o["t_dif"]
        for k in ("t_all", "u_int_before", "v_int_before", "w_int_before", "t_b"):
            assert gst[k] == pytest.approx(o[k], rel=SUM_TOL, abs=0.0), (l, k)
        for k in ("dif", "vel_dif", "u_dif", "v_dif", "w_dif"):
            assert gst[k] == pytest.approx(o[k], rel=1e-11, abs=1e-15), (l, k)


def test_monitor_no_difference(api):
    g = G.channel_grid(16, 9, 8)
    m = metrics(g)
    u, v, w, t, _ = fields(g, 4)
    st = api.MonitorState()
    api.monitor(g.nx, g.ny, g.nz, 1, dev(t), m["dx"], m["dy"], m["dz"], dev(t), dev(u),
                dev(v), dev(w), 1e-3, st)
    assert st["t_dif"] == 0.0 and st.loc == (0, 0, 0)


def o_convergence(g, u, re):
    r = ctypes.c_double(-1.0)
    oracle().o_check_convergence(P(g.y), g.yl, P(u), re, g.nx, g.ny, g.nz, ctypes.byref(r))
    return r.value


@pytest.mark.parametrize("n", SIZES)
def test_convergence_against_oracle(api, n):
    g = G.channel_grid(*n)
    m = metrics(g)
    u = np.random.default_rng(3).uniform(-2.0, 2.0, size=g.shape)
    got = api.check_convergence(m["y"], g.yl, dev(u), 5, 1e-3, 20.0, g.nx, g.ny, g.nz)
    assert got == pytest.approx(o_convergence(g, u, 20.0), rel=SUM_TOL, abs=0.0)


def test_convergence_poiseuille(api):
    re = 100.0
    g = G.set_c(16, 32, 8, 2.0 * math.pi, 2.0, math.pi)
    h = g.yl / g.ny
    yc = g.y[1: g.ny + 1]
    u = np.zeros(g.shape)
    u[1:-1, 1:-1, 1:-1] = (-re / 2.0 * yc * (yc - g.yl) + re * h * h / 8.0)[None, :, None]
    got = api.check_convergence(dev(g.y), g.yl, dev(u), 1, 1e-3, re, g.nx, g.ny, g.nz)
    assert got == pytest.approx(g.nz * g.ny * re * h * h / 8.0, rel=1e-12)
