"""The batched line-level solvers (CTRIDIAG, TDMA, cyclic,
implicit_solver, implicit_uy/vy/Tx/Ty) against the oracle's line
routines, which are pinned bit for bit against the reference's compiled
ones: every line of a batch equal to the oracle's solve of that line alone,
bit for bit.  Batches of 1 (the reference's own layout) and of several
hundred lines, on the stretched channel grid, with cellvalue in [0, 1],
random conductivities and heat capacities for the T systems, and periodic
line lengths of the FFT's 2^a 3^b form and otherwise.
"""
import ctypes

import numpy as np
import pytest

from oracle_bind import P, oracle
from vtkmakecompare_amd import grid as G

pytestmark = pytest.mark.gpu

RE, SC, DT, ETA = 2800.0, 0.71, 1e-3, 1e4


@pytest.fixture(scope="module")
def api():
    from vtkmakecompare_amd import api as A
    A.init()
    return A


def dev(A, a):
    return A.DeviceArray.from_host(np.ascontiguousarray(a, dtype=np.float64))


def col(a, q):
    return np.ascontiguousarray(a[..., q])


@pytest.mark.parametrize("nsys", [1, 300])
@pytest.mark.parametrize("n", [7, 48])
def test_tdma_and_cyclic(api, n, nsys):
    rng = np.random.default_rng(n + nsys)
    a = rng.uniform(-1, 0, size=(n + 2, nsys))
    c = rng.uniform(-1, 0, size=(n + 2, nsys))
    b = 3.0 + rng.uniform(0, 1, size=(n + 2, nsys))
    r = rng.normal(size=(n + 2, nsys))
    al, be = rng.uniform(-1, 1, size=nsys), rng.uniform(-1, 1, size=nsys)
    u = dev(api, np.zeros((n + 2, nsys)))
    api.tdma_batched(dev(api, a), dev(api, b), dev(api, c), dev(api, r), u, n, nsys)
    x = dev(api, np.zeros((n + 2, nsys)))
    api.cyclic_batched(dev(api, a), dev(api, b), dev(api, c), dev(api, al), dev(api, be),
                       dev(api, r), x, n, nsys)
    got_u, got_x = u.download(), x.download()
    o = oracle()
    for q in range(nsys):
        wu, wx = np.zeros(n + 2), np.zeros(n + 2)
        o.o_tdma(P(col(a, q)), P(col(b, q)), P(col(c, q)), P(col(r, q)), P(wu), n)
        o.o_cyclic(P(col(a, q)), P(col(b, q)), P(col(c, q)), al[q], be[q], P(col(r, q)),
                   P(wx), n)
        np.testing.assert_array_equal(got_u[:, q], wu)
        np.testing.assert_array_equal(got_x[:, q], wx)


@pytest.mark.parametrize("nsys", [1, 257])
def test_ctridiag(api, nsys):
    nn = 33
    rng = np.random.default_rng(nsys)
    a = np.empty((3, nn, nsys))
    a[0] = rng.uniform(-1, 0, size=(nn, nsys))
    a[2] = rng.uniform(-1, 0, size=(nn, nsys))
    a[1] = -3.0 - rng.uniform(0, 1, size=(nn, nsys))
    b = rng.normal(size=(nn, nsys, 2))
    da, db = dev(api, a), dev(api, b)
    api.ctridiag_batched(da, db, nn, nsys)
    got_a, got_b = da.download(), db.download()
    o = oracle()
    for q in range(nsys):
        wa = np.ascontiguousarray(a[:, :, q])  # A(NN,3) column-major = (3, NN) C-order
        wb = np.ascontiguousarray(b[:, q, :])
        o.o_ctridiag(P(wa), P(wb), nn)
        np.testing.assert_array_equal(got_b[:, q, :], wb)
        np.testing.assert_array_equal(got_a[:, :, q], wa)


@pytest.mark.parametrize("nsys", [1, 200])
def test_implicit_line_systems(api, nsys):
    g = G.channel_grid(24, 33, 18)
    rng = np.random.default_rng(17 + nsys)
    o = oracle()
    m = {k: dev(api, getattr(g, k)) for k in ("dx", "dx12", "dy", "dy12")}
    ft, fb = rng.normal(size=nsys), rng.normal(size=nsys)
    dft, dfb = dev(api, ft), dev(api, fb)

    # implicit_solver: SUX(0:nx+1), CUX(1:nx)
    nx = g.nx
    s = rng.normal(size=(nx + 2, nsys))
    cv = rng.uniform(0, 1, size=(nx, nsys)) * (rng.uniform(size=(nx, nsys)) < 0.3)
    d = dev(api, s)
    api.implicit_solver_batched(d, nx, m["dx"], m["dx12"], RE, DT, dft, dfb, ETA, dev(api, cv),
                                nsys)
    got = d.download()
    for q in range(nsys):
        w = col(s, q)
        o.o_implicit_solver(P(w), nx, P(g.dx), P(g.dx12), RE, DT, ft[q], fb[q], ETA,
                            P(col(cv, q)))
        np.testing.assert_array_equal(got[:, q], w)

    # implicit_uy: S(0:ny+1), C(1:ny); implicit_vy: S(0:ny), C(1:ny-1)
    ny = g.ny
    for name, ns, nc in (("uy", ny + 2, ny), ("vy", ny + 1, ny - 1)):
        s = rng.normal(size=(ns, nsys))
        cv = rng.uniform(0, 1, size=(nc, nsys))
        d = dev(api, s)
        getattr(api, f"implicit_{name}_batched")(d, ny, m["dy"], m["dy12"], RE, DT, dft, dfb,
                                                 ETA, dev(api, cv), nsys)
        got = d.download()
        for q in range(nsys):
            w = col(s, q)
            getattr(o, f"o_implicit_{name}")(P(w), ny, P(g.dy), P(g.dy12), RE, DT, ft[q], fb[q],
                                             ETA, P(col(cv, q)))
            np.testing.assert_array_equal(got[:, q], w, err_msg=name)

    # implicit_Tx / Ty: S, ll, rr, CU all (0:n+1)
    for name, n, dm, dm12, gm, gm12 in (("tx", nx, "dx", "dx12", g.dx, g.dx12),
                                        ("ty", ny, "dy", "dy12", g.dy, g.dy12)):
        s = rng.normal(size=(n + 2, nsys))
        ll = rng.uniform(0.5, 2.0, size=(n + 2, nsys))
        rr = rng.uniform(0.5, 2.0, size=(n + 2, nsys))
        cu = rng.uniform(0, 1, size=(n + 2, nsys))
        d = dev(api, s)
        if name == "tx":
            api.implicit_tx_batched(d, n, m[dm], m[dm12], RE, SC, DT, dev(api, ll),
                                    dev(api, rr), dft, dfb, ETA, dev(api, cu), nsys)
        else:
            api.implicit_ty_batched(d, n, m[dm], m[dm12], RE, SC, DT, dev(api, ll),
                                    dev(api, rr), dft, dfb, dev(api, cu), ETA, nsys)
        got = d.download()
        for q in range(nsys):
            w = col(s, q)
            if name == "tx":
                o.o_implicit_tx(P(w), n, P(gm), P(gm12), RE, SC, DT, P(col(ll, q)),
                                P(col(rr, q)), ft[q], fb[q], ETA, P(col(cu, q)))
            else:
                o.o_implicit_ty(P(w), n, P(gm), P(gm12), RE, SC, DT, P(col(ll, q)),
                                P(col(rr, q)), ft[q], fb[q], P(col(cu, q)), ETA)
            np.testing.assert_array_equal(got[:, q], w, err_msg=name)


def test_batched_refusals(api):
    d = dev(api, np.zeros((9, 4)))
    with pytest.raises(api.VKError):  # nsys 0
        api.tdma_batched(d, d, d, d, d, 7, 0)
    short = dev(api, np.zeros((9, 3)))
    with pytest.raises(api.VKError):  # an array too small for n and nsys
        api.tdma_batched(d, d, d, d, short, 7, 4)
