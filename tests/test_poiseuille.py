"""Laminar channel flow: the reference's own validation case.

CHECK_CONVERGENCE (nse_dominik.f:2875-2892) compares u(nx,j,k) with the
Poiseuille profile u = -Re/2 y (y - yl) that the constant forcing of
SET_CONVECT (the `+1.0D0*DT` of nse_dominik.f:785) drives in a channel of
height yl.  SET_CONVECTION_FOR_SCALAR adds the same unit source to T
(nse_dominik.f:879), whose diffusivity is 1/(Re Sc), so with rhoCp = thc
= 1 the temperature has the same profile with Re Sc in place of Re.

On the uniform grid of SET_C (h = yl/ny, centres y_j = (j - 1/2) h) the
discrete steady state is known in closed form:

  u*_j = -Re/2 y_j (y_j - yl) + Re h^2 / 8.

The second difference of a quadratic is exact, so the interior rows of
(u_{j+1} - 2 u_j + u_{j-1}) / h^2 = -Re hold for the parabola and for any
constant shift of it; the shift C is fixed by the wall reflection
u_0 = -u_1 of SET_BC: p(-h/2) + C = -(p(h/2) + C) gives C = Re h^2 / 8.

With v = w = 0 and u a function of y alone, every convective product
cancels exactly, the divergence is exactly zero (so MKP leaves u alone)
and one Crank-Nicolson step maps u* to itself: the AB2 forcing dt plus
dt/Re L u* vanishes (likewise T*).  A dropped source term, a wrong
diffusion coefficient, a wrong wall reflection or a wrong AB2 weight each
moves the fields off the profiles by O(dt) in one step, while rounding
alone leaves them within 1e-13 of it over ten steps.

The profile peaks at Re/2, and the convective terms are explicit (AB2),
so the case is run at Reynolds numbers where u dt/dx stays well below 1.
At Re 2800 the peak is 1400, and rounding noise in v and w grows by an
order of magnitude per step, in the oracle as in the library.
"""
import math

import numpy as np
import pytest

from oracle_bind import OracleRun
from vtkmakecompare_amd import grid as G


def discrete_profile(g, c):
    """-c/2 y (y - yl) + c h^2 / 8 at the cell centres j = 1..ny."""
    h = g.yl / g.ny
    yc = g.y[1: g.ny + 1]
    return -c / 2.0 * yc * (yc - g.yl) + c * h * h / 8.0


def poiseuille_case(nx, ny, nz, re, sc=0.71):
    """Uniform SET_C grid of the channel [0, 2 pi] x [0, 2] x [0, pi] with
    u = u*(Re), T = T*(Re Sc), v = w = 0, no solid."""
    g = G.set_c(nx, ny, nz, 2.0 * math.pi, 2.0, math.pi)
    cfg = G.Config(f"poiseuille{nx}x{ny}x{nz}", nx, ny, nz, re, sc, 1e-3, 1e4, 10,
                   0x5EED00C0, 0.0, [])
    st = {k: np.zeros(g.shape) for k in G.FIELDS}
    prof = discrete_profile(g, re)
    st["u"][1:-1, 1:-1, 1:-1] = prof[None, :, None]
    st["t"][1:-1, 1:-1, 1:-1] = discrete_profile(g, re * sc)[None, :, None]
    st["rhocp"][...] = 1.0
    st["thc"][...] = 1.0
    for a, b in (("u_pre", "u"), ("v_pre", "v"), ("w_pre", "w"), ("t_pre", "t")):
        st[a][...] = st[b]
    st["phi0"] = np.zeros(g.shape)
    st["phi"] = np.zeros(g.shape)
    st["cellvalue"] = np.zeros((4, nz, ny, nx))
    return cfg, g, st, prof


def analytic_error_sum(u, g, re):
    """CHECK_CONVERGENCE's sum over the plane i = nx of |u - u_analytic|."""
    yc = g.y[1: g.ny + 1]
    ana = -re / 2.0 * yc * (yc - g.yl)
    return float(np.sum(np.abs(u[1:-1, 1:-1, g.nx] - ana[None, :])))


def profile_drift(u, prof):
    """max |u - u*| / max |u*| over the interior."""
    return float(np.max(np.abs(u[1:-1, 1:-1, 1:-1] - prof[None, :, None])) / np.max(prof))


@pytest.mark.parametrize("re", [10.0, 100.0])
def test_oracle_keeps_poiseuille(re):
    """The oracle's step has u* and T* as fixed points (l = 1 and the AB2
    steps after it), and CHECK_CONVERGENCE's error sum is nz ny Re h^2 / 8."""
    cfg, g, st, prof = poiseuille_case(8, 16, 8, re)
    ora = OracleRun(cfg, g, st)
    ora.step(10)
    assert profile_drift(ora["u"], prof) < 1e-13
    assert profile_drift(ora["t"], discrete_profile(g, re * cfg.sc)) < 1e-13
    for k in ("v", "w"):
        assert np.max(np.abs(ora[k][1:-1, 1:-1, 1:-1])) < 1e-12 * np.max(prof), k
    h = g.yl / g.ny
    expect = g.nz * g.ny * re * h * h / 8.0
    assert analytic_error_sum(ora["u"], g, re) == pytest.approx(expect, rel=1e-10)


def test_oracle_poiseuille_is_sharp():
    """The fixed point is sharp: the continuum parabola (the shift C left
    out) drifts by O(dt) in one step, and so does u* at a Reynolds number
    other than the one it was built for."""
    cfg, g, st, prof = poiseuille_case(8, 16, 8, 100.0)
    h = g.yl / g.ny
    st["u"][1:-1, 1:-1, 1:-1] -= 100.0 * h * h / 8.0
    ora = OracleRun(cfg, g, st)
    ora.step(1)
    assert profile_drift(ora["u"], prof - 100.0 * h * h / 8.0) > 1e-7
    cfg, g, st, prof = poiseuille_case(8, 16, 8, 100.0)
    cfg.re = 110.0
    ora = OracleRun(cfg, g, st)
    ora.step(1)
    assert profile_drift(ora["u"], prof) > 1e-7


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol,tol_t,sum_rel", [("fp64", 1e-13, 1e-13, 1e-9),
                                                         ("mixed", 1e-6, 1e-5, 1e-2)])
@pytest.mark.parametrize("n", [(16, 32, 16), (64, 65, 64)])
def test_gpu_keeps_poiseuille(n, precision, tol, tol_t, sum_rel):
    """The library's fused step (both forms) keeps u* and T* for ten steps,
    and v and w stay at zero.  The mixed form's bars are its velocity and
    temperature bars against the oracle."""
    from vtkmakecompare_amd import api as A
    A.init()
    re = 100.0
    cfg, g, st, prof = poiseuille_case(*n, re)
    run = A.ChannelRun(cfg, g, st, cellvalue=st["cellvalue"], precision=precision)
    run.step(10)
    u = run.get("u")
    assert profile_drift(u, prof) < tol
    assert profile_drift(run.get("t"), discrete_profile(g, re * cfg.sc)) < tol_t
    for k in ("v", "w"):
        assert np.max(np.abs(run.get(k)[1:-1, 1:-1, 1:-1])) < tol * np.max(prof), k
    h = g.yl / g.ny
    expect = g.nz * g.ny * re * h * h / 8.0
    assert analytic_error_sum(u, g, re) == pytest.approx(expect, rel=sum_rel)
