"""Pin the C oracle against the reference routines compiled from the
reference's own Fortran sources (oracle/_ref/libref.so).

The two are compiled without fused multiply-add and perform the same
operations in the same order, so most comparisons are bit-for-bit; the
complex/real divisions of CTRIDIAG/SP0B may differ in the last bit
(Fortran mixed-mode division goes through complex division), so those use
a 1e-14 normwise bound.  Skipped where the reference was not built (the
GPU box has no reference sources).
"""
import os

import numpy as np
import pytest

from oracle_bind import D, I, P, oracle, ref
from vtkmakecompare_amd import grid as G

pytestmark = pytest.mark.skipif(ref() is None, reason="reference library not built")

SIZES = [(8, 9, 8), (12, 7, 6)]


def rfield(rng, g, lo=-1.0, hi=1.0):
    return rng.uniform(lo, hi, size=g.shape)


def rcell(rng, g):
    cv = np.zeros((4, g.nz, g.ny, g.nx))
    cv[0] = rng.uniform(0.0, 1.0, size=(g.nz, g.ny, g.nx))
    cv[1:] = rng.uniform(0.0, 1.0, size=(3, g.nz, g.ny, g.nx))
    return cv


def grid_args(g):
    return [P(g.dx), P(g.dy), P(g.dz), P(g.dx12), P(g.dy12), P(g.dz12)]


def mk_grid(n):
    return G.channel_grid(*n)


def exact(a, b):
    np.testing.assert_array_equal(a, b)


def close(a, b, tol=1e-14):
    scale = max(np.max(np.abs(b)), 1e-300)
    assert np.max(np.abs(a - b)) <= tol * scale, np.max(np.abs(a - b)) / scale


@pytest.mark.parametrize("n", SIZES)
def test_set_c(n):
    nx, ny, nz = n
    arrs_r = [np.zeros(k + 2) for k in (nx, nx, ny, ny, nz, nz, nx, nx, ny, ny, nz, nz)]
    arrs_o = [np.zeros(k + 2) for k in (nx, nx, ny, ny, nz, nz, nx, nx, ny, ny, nz, nz)]
    xl, yl, zl = 2.5e-1, 4.0, 3.3
    ref().set_c_(I(nx), I(ny), I(nz), *[P(a) for a in arrs_r], D(xl), D(yl), D(zl))
    oracle().o_set_c(nx, ny, nz, *[P(a) for a in arrs_o], xl, yl, zl)
    g = G.set_c(nx, ny, nz, xl, yl, zl)
    host = [g.x, g.x12, g.y, g.y12, g.z, g.z12, g.dx, g.dx12, g.dy, g.dy12, g.dz, g.dz12]
    for a, b, c in zip(arrs_r, arrs_o, host):
        exact(a, b)
        exact(a, c)


@pytest.mark.parametrize("n", SIZES)
def test_set_bc(n):
    g = mk_grid(n)
    rng = np.random.default_rng(1)
    fr = [rfield(rng, g) for _ in range(5)]
    fo = [a.copy() for a in fr]
    ref().set_bc_(P(fr[0]), P(fr[1]), P(fr[2]), P(fr[3]), I(g.nx), I(g.ny), I(g.nz), P(fr[4]))
    oracle().o_set_bc(P(fo[0]), P(fo[1]), P(fo[2]), P(fo[3]), g.nx, g.ny, g.nz, P(fo[4]))
    for a, b in zip(fr, fo):
        exact(a, b)


@pytest.mark.parametrize("n", SIZES)
def test_stencils(n):
    g = mk_grid(n)
    rng = np.random.default_rng(2)
    u, v, w, t = (rfield(rng, g) for _ in range(4))
    rc = rfield(rng, g, 0.5, 2.0)
    th = rfield(rng, g, 0.5, 2.0)
    dt, re, sc = 1e-3, 2800.0, 0.71
    outs_r = [np.zeros(g.shape) for _ in range(8)]
    outs_o = [np.zeros(g.shape) for _ in range(8)]
    R, O = ref(), oracle()
    R.set_convect_(P(u), P(outs_r[0]), P(v), P(outs_r[1]), P(w), P(outs_r[2]), D(dt),
                   I(g.nx), I(g.ny), I(g.nz), *grid_args(g))
    O.o_set_convect(P(u), P(outs_o[0]), P(v), P(outs_o[1]), P(w), P(outs_o[2]), dt,
                    g.nx, g.ny, g.nz, *grid_args(g))
    R.set_convection_for_scalar_(P(u), P(v), P(w), D(dt), I(g.nx), I(g.ny), I(g.nz),
                                 *grid_args(g), P(t), P(outs_r[3]), P(rc))
    O.o_set_convection_for_scalar(P(u), P(v), P(w), dt, g.nx, g.ny, g.nz, *grid_args(g),
                                  P(t), P(outs_o[3]), P(rc))
    R.set_conduction_(P(u), P(outs_r[4]), P(v), P(outs_r[5]), P(w), P(outs_r[6]), D(re),
                      D(dt), I(g.nx), I(g.ny), I(g.nz), *grid_args(g))
    O.o_set_conduction(P(u), P(outs_o[4]), P(v), P(outs_o[5]), P(w), P(outs_o[6]), re, dt,
                       g.nx, g.ny, g.nz, *grid_args(g))
    R.set_diffusion_for_scalar_(D(re), D(dt), I(g.nx), I(g.ny), I(g.nz), *grid_args(g),
                                P(t), P(outs_r[7]), D(sc), P(th), P(rc))
    O.o_set_diffusion_for_scalar(re, dt, g.nx, g.ny, g.nz, *grid_args(g), P(t),
                                 P(outs_o[7]), sc, P(th), P(rc))
    for a, b in zip(outs_r, outs_o):
        exact(a, b)


@pytest.mark.parametrize("l", [1, 2])
def test_immersed_and_explicit(l):
    g = mk_grid((8, 9, 8))
    rng = np.random.default_rng(3 + l)
    cv_r = rcell(rng, g)
    cv_o = cv_r.copy()
    uta, vta, wta = (rfield(rng, g) for _ in range(3))
    eta, dt = 1e4, 1e-3
    lr = [np.zeros(g.shape) for _ in range(4)]
    lo = [np.zeros(g.shape) for _ in range(4)]
    cons = np.zeros(1)
    ref().set_immersed_boundary_(D(eta), I(g.nx), I(g.ny), I(g.nz), D(dt), P(lr[0]), P(lr[1]),
                                 P(lr[2]), P(g.dx), P(g.dy), P(g.dz), P(lr[3]), P(cv_r),
                                 P(uta), P(vta), P(wta), I(l), P(cons))
    oracle().o_set_immersed_boundary(eta, g.nx, g.ny, g.nz, dt, P(lo[0]), P(lo[1]), P(lo[2]),
                                     P(g.dx), P(g.dy), P(g.dz), P(lo[3]), P(cv_o), P(uta),
                                     P(vta), P(wta), l, P(cons))
    exact(cv_r, cv_o)
    for a, b in zip(lr, lo):
        exact(a, b)
    # SET_EXPLICIT and the temperature variant
    ins = [rfield(rng, g) for _ in range(15)]
    ins_o = [a.copy() for a in ins]
    ref().set_explicit_(*[P(a) for a in ins], I(g.nx), I(g.ny), I(g.nz), I(l))
    oracle().o_set_explicit(*[P(a) for a in ins_o], g.nx, g.ny, g.nz, l)
    for a, b in zip(ins, ins_o):
        exact(a, b)
    tin = [rfield(rng, g) for _ in range(5)]
    tin_o = [a.copy() for a in tin]
    ref().set_explicit_for_tempreture_(*[P(a) for a in tin], I(g.nx), I(g.ny), I(g.nz), I(l))
    oracle().o_set_explicit_for_tempreture(*[P(a) for a in tin_o], g.nx, g.ny, g.nz, l)
    for a, b in zip(tin, tin_o):
        exact(a, b)


@pytest.mark.parametrize("ny", [5, 17, 64])
def test_line_solvers(ny):
    g = G.channel_grid(8, ny, 8)
    rng = np.random.default_rng(ny)
    re, sc, dt, eta = 2800.0, 0.71, 1e-3, 1e4
    cv = rng.uniform(0, 1, ny + 2)
    # implicit_uy: SUY(0:ny+1), CUY(1:ny)
    s_r = rng.uniform(-1, 1, ny + 2)
    s_o = s_r.copy()
    cuy = np.ascontiguousarray(cv[1:ny + 1])
    ref().implicit_uy_(P(s_r), I(ny), P(g.dy), P(g.dy12), D(re), D(dt), D(0.0), D(0.0),
                       D(eta), P(cuy))
    oracle().o_implicit_uy(P(s_o), ny, P(g.dy), P(g.dy12), re, dt, 0.0, 0.0, eta, P(cuy))
    exact(s_r, s_o)
    # implicit_vy
    s_r = rng.uniform(-1, 1, ny + 2)
    s_o = s_r.copy()
    ref().implicit_vy_(P(s_r), I(ny), P(g.dy), P(g.dy12), D(re), D(dt), D(0.0), D(0.0),
                       D(eta), P(cuy))
    oracle().o_implicit_vy(P(s_o), ny, P(g.dy), P(g.dy12), re, dt, 0.0, 0.0, eta, P(cuy))
    exact(s_r, s_o)
    # implicit_Ty with defined ghosts
    lly = rng.uniform(0.5, 2.0, ny + 2)
    rry = rng.uniform(0.5, 2.0, ny + 2)
    s_r = rng.uniform(-1, 1, ny + 2)
    s_o = s_r.copy()
    ref().implicit_ty_(P(s_r), I(ny), P(g.dy), P(g.dy12), D(re), D(sc), D(dt), P(lly), P(rry),
                       D(0.0), D(0.0), P(cv), D(eta))
    oracle().o_implicit_ty(P(s_o), ny, P(g.dy), P(g.dy12), re, sc, dt, P(lly), P(rry),
                           0.0, 0.0, P(cv), eta)
    exact(s_r, s_o)
    # periodic: implicit_solver and implicit_Tx on the uniform x axis
    n = ny
    gx = G.channel_grid(n, 5, 4)
    s_r = rng.uniform(-1, 1, n + 2)
    s_o = s_r.copy()
    cux = np.ascontiguousarray(rng.uniform(0, 1, n))
    ref().implicit_solver_(P(s_r), I(n), P(gx.dx), P(gx.dx12), D(re), D(dt), D(0.0), D(0.0),
                           D(eta), P(cux))
    oracle().o_implicit_solver(P(s_o), n, P(gx.dx), P(gx.dx12), re, dt, 0.0, 0.0, eta, P(cux))
    exact(s_r, s_o)
    llx = rng.uniform(0.5, 2.0, n + 2)
    rrx = rng.uniform(0.5, 2.0, n + 2)
    cux2 = rng.uniform(0, 1, n + 2)
    s_r = rng.uniform(-1, 1, n + 2)
    s_o = s_r.copy()
    ref().implicit_tx_(P(s_r), I(n), P(gx.dx), P(gx.dx12), D(re), D(sc), D(dt), P(llx), P(rrx),
                       D(0.0), D(0.0), D(eta), P(cux2))
    oracle().o_implicit_tx(P(s_o), n, P(gx.dx), P(gx.dx12), re, sc, dt, P(llx), P(rrx),
                           0.0, 0.0, eta, P(cux2))
    exact(s_r, s_o)


def test_solve_u_temp():
    g = mk_grid((8, 9, 6))
    rng = np.random.default_rng(7)
    cv = rcell(rng, g)
    a = [rfield(rng, g) for _ in range(3)]
    old = [rfield(rng, g) for _ in range(3)]
    ao = [x.copy() for x in a]
    re, dt, eta = 2800.0, 1e-3, 1e4
    ref().solve_u_temp_(P(a[0]), P(a[1]), P(a[2]), I(g.nx), I(g.ny), I(g.nz), *grid_args(g),
                        D(re), D(dt), D(eta), P(cv), P(old[0]), P(old[1]), P(old[2]))
    oracle().o_solve_u_temp(P(ao[0]), P(ao[1]), P(ao[2]), g.nx, g.ny, g.nz, *grid_args(g),
                            re, dt, eta, P(cv), P(old[0]), P(old[1]), P(old[2]))
    for x, y in zip(a, ao):
        exact(x, y)


def test_spectral_solvers():
    nx, ny, nz = 8, 9, 6
    g = G.channel_grid(nx, ny, nz)
    rng = np.random.default_rng(8)
    nh = nx // 2
    kxz_r = np.zeros((nz + 1, nh + 1))
    ac = np.zeros((3, ny))
    kxz = np.zeros((nz + 1, nh + 1))
    oracle().o_mkp_coefficients(nx, ny, nz, P(g.dx), P(g.dy), P(g.dz), P(g.dy12),
                                g.xl, g.zl, P(kxz), P(ac))
    phit = rng.normal(size=(nz + 2, ny + 2, nh + 1, 2))
    pr, po = phit.copy(), phit.copy()
    ref().spn_(I(nx), I(ny), I(nz), P(ac), P(kxz), P(pr))
    oracle().o_spn(nx, ny, nz, P(ac), P(kxz), P(po))
    close(po, pr)
    ref().sp0b_(I(nx), I(ny), I(nz), P(ac), P(pr))
    oracle().o_sp0b(nx, ny, nz, P(ac), P(po))
    close(po, pr)
    del kxz_r
    # CTRIDIAG on its own
    a = np.ascontiguousarray(rng.uniform(-1, 1, (3, 11)))
    a[1] += 4.0
    b = rng.normal(size=(11, 2))
    ar, br, ao, bo = a.copy(), b.copy(), a.copy(), b.copy()
    ref().ctridiag_(P(ar), P(br), I(11))
    oracle().o_ctridiag(P(ao), P(bo), 11)
    close(bo, br)
    # PHIBC
    phi = rfield(rng, g)
    phr, pho = phi.copy(), phi.copy()
    ref().phibc_(I(nx), I(ny), I(nz), P(phr))
    oracle().o_phibc(nx, ny, nz, P(pho))
    exact(phr, pho)


def test_set_levelset(tmp_path):
    g = G.channel_grid(8, 9, 8)
    phi0 = G.sphere_levelset(g, [(np.pi, 0.1, 0.5 * np.pi, 0.9)])
    cv_r = np.zeros((4, g.nz, g.ny, g.nx))
    phi_r = np.zeros(g.shape)
    cwd = os.getcwd()
    os.chdir(tmp_path)  # the reference writes two VTK files
    try:
        ref().set_levelset_(I(g.nx), I(g.ny), I(g.nz), P(g.dx), P(g.dx12), P(g.dy12),
                            P(g.dz12), P(phi0), P(phi_r), P(g.x), P(g.y), P(g.z), P(cv_r))
    finally:
        os.chdir(cwd)
    from oracle_bind import oracle_levelset
    phi_o, cv_o = oracle_levelset(g, phi0, 2001)
    exact(phi_r, phi_o)
    exact(cv_r[0], cv_o[0])


def test_adjoint_kernels():
    g = mk_grid((8, 9, 6))
    rng = np.random.default_rng(9)
    u, v, w, uf, vf, wf, t = (rfield(rng, g) for _ in range(7))
    rc = rfield(rng, g, 0.5, 2.0)
    re, dt = 2800.0, 1e-3
    outr = [np.zeros(g.shape) for _ in range(4)]
    outo = [np.zeros(g.shape) for _ in range(4)]
    ref().set_adj_convect_(P(u), P(outr[0]), P(v), P(outr[1]), P(w), P(outr[2]), D(re), D(dt),
                           I(g.nx), I(g.ny), I(g.nz), *grid_args(g), P(uf), P(vf), P(wf))
    oracle().o_set_adj_convect(P(u), P(outo[0]), P(v), P(outo[1]), P(w), P(outo[2]), re, dt,
                               g.nx, g.ny, g.nz, *grid_args(g), P(uf), P(vf), P(wf))
    ref().set_adj_convection_for_scalar_(P(u), P(v), P(w), D(re), D(dt), I(g.nx), I(g.ny),
                                         I(g.nz), *grid_args(g), P(t), P(outr[3]), P(rc),
                                         P(uf), P(vf), P(wf))
    oracle().o_set_adj_convection_for_scalar(P(u), P(v), P(w), re, dt, g.nx, g.ny, g.nz,
                                             *grid_args(g), P(t), P(outo[3]), P(rc),
                                             P(uf), P(vf), P(wf))
    for a, b in zip(outr, outo):
        exact(a, b)


# ---------------------------------------------- iterative Poisson (dead)

def _poisson_case(n, seed):
    g = mk_grid(n)
    rng = np.random.default_rng(seed)
    ut, vt, wt = (rfield(rng, g) for _ in range(3))
    p = rfield(rng, g, -0.1, 0.1)
    return g, ut, vt, wt, p


@pytest.mark.parametrize("n", [(8, 7, 6), (12, 9, 12)])
def test_linebyline_bit_exact(n):
    """nse_dominik.f:2136-2267: three iterations of the line Gauss-Seidel
    sweep from the same p (grids with ny, nz <= nx + 1, where the
    reference's fpy and fpz hold their lines)."""
    g, ut, vt, wt, p = _poisson_case(n, 31)
    a, b = p.copy(), p.copy()
    dt = 1e-3
    for _ in range(3):
        oracle().o_linebyline(g.nx, g.ny, g.nz, P(ut), P(vt), P(wt), *grid_args(g), dt, P(a))
        ref().linebyline_(I(g.nx), I(g.ny), I(g.nz), P(ut), P(vt), P(wt), *grid_args(g),
                          D(dt), P(b))
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("uniform", [True, False])
def test_solve_p_bit_exact(uniform):
    """nse_dominik.f:1955-2021: linebyline until the 1e-5 residual; p and
    the four residual statistics.  On the uniform grid the velocity has no
    flux through the boundary (a compatible right-hand side for the
    Neumann problem) and the iteration converges; on the stretched grid
    every line solve uses its last row's metrics, the fixed point is not
    the discrete solution and the reference runs all 100000 iterations."""
    import ctypes
    g, ut, vt, wt, p = _poisson_case((10, 8, 8), 37)
    if uniform:
        g = G.set_c(10, 8, 8, 2.0, 2.0, 2.0)
        ut[:, :, 0] = ut[:, :, 10] = 0.0
        vt[:, 0, :] = vt[:, 8, :] = 0.0
        wt[0] = wt[8] = 0.0
    ut *= 1e-3
    vt *= 1e-3
    wt *= 1e-3
    dt = 1e-3
    a, b = p.copy(), p.copy()
    ro = [ctypes.c_double(0.0) for _ in range(4)]
    rr = [ctypes.c_double(0.0) for _ in range(4)]
    it = ctypes.c_int(0)
    oracle().o_solve_p(P(a), P(ut), P(vt), P(wt), dt, *grid_args(g), g.nx, g.ny, g.nz,
                       *[ctypes.byref(x) for x in ro], 100000, ctypes.byref(it))
    ref().solve_p_(P(b), P(ut), P(vt), P(wt), D(dt), *grid_args(g), I(g.nx), I(g.ny),
                   I(g.nz), *[ctypes.byref(x) for x in rr])
    assert it.value >= 1 and (it.value < 100000) == uniform, it.value
    np.testing.assert_array_equal(a, b)
    assert [x.value for x in ro] == [x.value for x in rr]
