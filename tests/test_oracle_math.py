"""Pins for the oracle functions the reference build cannot run:
the FFT and the P2S/S2P conventions (the reference needs FFTW, absent
here), MKP as a whole (it calls P2S/S2P) and set_adj_NS_scalar (it reads an
uninitialised local).  Each check is independent of the oracle's own
formulas: numpy's pocketfft, a dense least-squares Poisson solve, the
post-projection divergence, and closed forms.
"""
import math

import numpy as np
import pytest

from oracle_bind import P, oracle
from vtkmakecompare_amd import grid as G


def nrel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 8, 12, 16, 24, 48, 64, 96, 128, 192, 384, 7])
def test_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    buf = np.stack([x.real, x.imag], axis=-1).copy()
    oracle().o_fft(P(buf), n, 1, -1)
    got = buf[:, 0] + 1j * buf[:, 1]
    assert nrel(got, np.fft.fft(x)) < 1e-13
    buf = np.stack([x.real, x.imag], axis=-1).copy()
    oracle().o_fft(P(buf), n, 1, +1)
    got = buf[:, 0] + 1j * buf[:, 1]
    assert nrel(got, np.fft.ifft(x) * n) < 1e-13


def _wrapped_field(rng, nx, ny, nz):
    fp = rng.normal(size=(nz + 2, ny + 2, nx + 2))
    return fp


@pytest.mark.parametrize("n", [(8, 5, 6), (12, 3, 24), (16, 4, 8)])
def test_p2s_s2p_conventions(n):
    nx, ny, nz = n
    rng = np.random.default_rng(nx * nz)
    fp = _wrapped_field(rng, nx, ny, nz)
    nh = nx // 2
    fpt = np.zeros((nz + 2, ny + 2, nh + 1, 2))
    oracle().o_p2s(P(fp), P(fpt), nx, ny, nz)
    for j in range(1, ny + 1):
        plane = fp[0:nz, j, 0:nx]          # samples I,K = 0..N-1 (index 0 is the halo)
        ref = np.fft.rfft2(plane) / (nx * nz)
        got = fpt[0:nz, j, :, 0] + 1j * fpt[0:nz, j, :, 1]
        assert nrel(got, ref) < 1e-13
    # a single mode maps to a single spike
    one = np.zeros_like(fp)
    kk, ii = 2, 1
    for k in range(nz):
        for i in range(nx):
            one[k, 1, i] = np.cos(2 * np.pi * (ii * i / nx + kk * k / nz))
    spk = np.zeros_like(fpt)
    oracle().o_p2s(P(one), P(spk), nx, ny, nz)
    mag = np.hypot(spk[0:nz, 1, :, 0], spk[0:nz, 1, :, 1])
    assert abs(mag[kk, ii] - 0.5) < 1e-14 and np.sum(mag > 1e-12) == 1
    # S2P: unnormalised inverse with wrap copies, and the round trip
    back = np.zeros_like(fp)
    oracle().o_s2p(P(fpt), P(back), nx, ny, nz)
    for j in range(1, ny + 1):
        spec = fpt[0:nz, j, :, 0] + 1j * fpt[0:nz, j, :, 1]
        ref = np.fft.irfft2(spec, s=(nz, nx)) * (nx * nz)
        assert nrel(back[0:nz, j, 0:nx], ref) < 1e-13
        assert nrel(back[0:nz, j, 0:nx], fp[0:nz, j, 0:nx]) < 1e-13
        np.testing.assert_array_equal(back[0:nz + 1, j, nx], back[0:nz + 1, j, 0])
        np.testing.assert_array_equal(back[nz, j, 0:nx + 1], back[0, j, 0:nx + 1])


def _dense_poisson(g):
    """The operator MKP inverts: periodic 3-point second differences in x
    and z, the flux-form y operator with Neumann walls (mkp.f:80-84,
    247-250), on the interior points in (k, j, i) order."""
    nx, ny, nz = g.nx, g.ny, g.nz
    N = nx * ny * nz
    A = np.zeros((N, N))

    def idx(i, j, k):
        return ((k % nz) * ny + j) * nx + (i % nx)

    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                r = idx(i, j, k)
                dx2 = g.dx[1] ** 2
                dz2 = g.dz[1] ** 2
                for di in (-1, 1):
                    A[r, idx(i + di, j, k)] += 1.0 / dx2
                A[r, r] -= 2.0 / dx2
                for dk in (-1, 1):
                    A[r, idx(i, j, k + dk)] += 1.0 / dz2
                A[r, r] -= 2.0 / dz2
                J = j + 1
                a1 = 1.0 / g.dy12[J - 1] / g.dy[J]
                a3 = 1.0 / g.dy12[J] / g.dy[J]
                A[r, r] -= a1 + a3
                if J > 1:
                    A[r, idx(i, j - 1, k)] += a1
                else:
                    A[r, r] += a1
                if J < ny:
                    A[r, idx(i, j + 1, k)] += a3
                else:
                    A[r, r] += a3
    return A


def _mkp_inputs(g, rng):
    ut = [rng.normal(size=g.shape) for _ in range(3)]
    p = rng.normal(size=g.shape)
    t = np.zeros(g.shape)
    oracle().o_set_bc(P(ut[0]), P(ut[1]), P(ut[2]), P(p), g.nx, g.ny, g.nz, P(t))
    return ut, p


def test_mkp_against_dense_solve():
    g = G.channel_grid(8, 9, 6)
    rng = np.random.default_rng(11)
    ut, p = _mkp_inputs(g, rng)
    outs = {n: np.zeros(g.shape) for n in
            ("u_ta", "v_ta", "w_ta", "u_ep2", "v_ep2", "w_ep2", "divu", "u", "v", "w")}
    dt = 1e-3
    oracle().o_mkp(P(p), P(ut[0]), P(ut[1]), P(ut[2]), dt, P(g.dx), P(g.dy), P(g.dz),
                   P(g.dx12), P(g.dy12), P(g.dz12), 2800.0, g.nx, g.ny, g.nz,
                   *[P(outs[n]) for n in ("u_ta", "v_ta", "w_ta", "u_ep2", "v_ep2",
                                          "w_ep2", "divu", "u", "v", "w")], g.xl, g.zl)
    u, v, w = ut
    rhs = ((u[1:-1, 1:-1, 1:-1] - u[1:-1, 1:-1, :-2]) / g.dx[None, None, 1:-1] +
           (v[1:-1, 1:-1, 1:-1] - v[1:-1, :-2, 1:-1]) / g.dy[None, 1:-1, None] +
           (w[1:-1, 1:-1, 1:-1] - w[:-2, 1:-1, 1:-1]) / g.dz[1:-1, None, None]) / dt
    A = _dense_poisson(g)
    phi = np.linalg.lstsq(A, rhs.ravel(), rcond=None)[0].reshape(g.nz, g.ny, g.nx)
    phi -= phi[:, 0, :].mean()           # SP0B gauge: zero plane mean at j = 1
    got = p[1:-1, 1:-1, 1:-1]
    assert abs(got[:, 0, :].mean()) < 1e-9 * np.max(np.abs(got))
    assert nrel(got, phi) < 1e-10
    # the projection leaves a divergence-free field
    div = outs["divu"][1:-1, 1:-1, 1:-1]
    assert np.max(np.abs(div)) < 1e-10 * np.max(np.abs(rhs * dt))
    # u_ta = dt * grad(phi) at the faces, u = u_temp - u_ta
    np.testing.assert_allclose(outs["u"][1:-1, 1:-1, 1:-1],
                               ut[0][1:-1, 1:-1, 1:-1] - outs["u_ta"][1:-1, 1:-1, 1:-1],
                               rtol=0, atol=1e-12)


def test_adj_ns_scalar_closed_form():
    g = G.channel_grid(8, 9, 6)
    nx, ny, nz = g.nx, g.ny, g.nz
    a, b, c = 0.7, -1.3, 2.1
    X = g.x[None, None, :]
    Y = g.y[None, :, None]
    Z = g.z[:, None, None]
    t = np.ascontiguousarray(np.broadcast_to(a * X + b * Y + c * Z, g.shape))
    tf = np.full(g.shape, 3.0)
    us, vs, ws = (np.zeros(g.shape) for _ in range(3))
    dummy = np.zeros(g.shape)
    dt = 1e-3
    oracle().o_set_adj_ns_scalar(dt, nx, ny, nz, P(g.dx12), P(g.dy12), P(g.dz12), P(t),
                                 P(us), P(vs), P(ws), P(dummy), P(dummy), P(dummy), P(tf))
    # centred difference over dx12 on a uniform x axis gives the slope
    np.testing.assert_allclose(us[1:-1, 1:-1, 1:-1], dt * 3.0 * a, rtol=1e-12)
    np.testing.assert_allclose(ws[1:-1, 1:-1, 1:-1], dt * 3.0 * c, rtol=1e-12)
    # on the stretched y axis the reference divides the two-cell
    # difference by 2*dy12(j) (adjoint.f:204)
    jj = np.arange(1, ny + 1)
    expect = dt * 3.0 * b * (g.y[jj + 1] - g.y[jj - 1]) / g.dy12[jj] / 2.0
    np.testing.assert_allclose(vs[1:-1, 1:-1, 1:-1],
                               np.broadcast_to(expect[None, :, None], (nz, ny, nx)),
                               rtol=1e-12)


# ------------------------------------------------ analytic shape union

def _shapes_grid():
    return G.channel_grid(24, 17, 20)


def test_shapes_sphere_is_signed_distance():
    """An axis-aligned sphere gives r - |d| (the exact signed distance,
    positive inside), with the minimum image in x and z: a centre at
    x = 0.1 is seen across the periodic boundary at x = xl - 0.1."""
    from oracle_bind import oracle_shapes
    g = _shapes_grid()
    spheres = [(0.1, 0.2, 0.5 * g.zl, 0.45), (0.6 * g.xl, -0.4, 0.05, 0.3)]
    recs = [(cx, cy, cz, r, r, r, 1.0, 0.0, 0.0, 0.0) for cx, cy, cz, r in spheres]
    got = oracle_shapes(g, recs)
    want = G.sphere_levelset(g, spheres)
    assert np.max(np.abs(got - want)) < 1e-15
    assert np.all(got[0] == 0.0) and np.all(got[:, :, -1] == 0.0)


def test_shapes_rotated_ellipsoids_against_scipy_rotation():
    """Rotated ellipsoids against body coordinates formed with scipy's
    rotation of the same quaternion (scalar-last there): catches a
    transposed rotation, a wrong quaternion order or a wrong axis."""
    from scipy.spatial.transform import Rotation

    from oracle_bind import oracle_shapes
    g = _shapes_grid()
    recs = G.random_ellipsoids(0x5EED0FFF, 5, g.xl, g.zl, 0.2, 0.7)
    got = oracle_shapes(g, recs)
    X = g.x[1:-1][None, None, :]
    Y = g.y[1:-1][None, :, None]
    Z = g.z[1:-1][:, None, None]
    best = None
    for rec in recs:
        c, ax, q = rec[:3], rec[3:6], rec[6:]
        R = Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()
        dx = X - c[0]
        dx = np.where(dx > 0.5 * g.xl, dx - g.xl, np.where(dx < -0.5 * g.xl, dx + g.xl, dx))
        dz = Z - c[2]
        dz = np.where(dz > 0.5 * g.zl, dz - g.zl, np.where(dz < -0.5 * g.zl, dz + g.zl, dz))
        d = np.stack(np.broadcast_arrays(dx, Y - c[1], dz), axis=-1)
        body = d @ R  # R^T d for each point
        v = ax.min() * (1.0 - np.sqrt(np.sum((body / ax) ** 2, axis=-1)))
        best = v if best is None else np.maximum(best, v)
    assert np.max(np.abs(got[1:-1, 1:-1, 1:-1] - best)) < 1e-14


def test_shapes_surface_and_sign():
    """A cell centre on the surface of a rotated ellipsoid gives 0: the
    quaternion of a quarter turn about z maps the body x axis onto world
    y, so the point a along world y from the centre is on the surface."""
    from oracle_bind import oracle_shapes
    g = G.set_c(16, 16, 16, 4.0, 4.0, 4.0)
    i, j, k = 8, 8, 8
    a = g.y[j + 3] - g.y[j]
    h = math.sqrt(0.5)
    rec = (g.x[i], g.y[j], g.z[k], a, 0.5 * a, 0.4 * a, h, 0.0, 0.0, h)
    phi = oracle_shapes(g, [rec])
    assert abs(phi[k, j + 3, i]) < 1e-15
    assert phi[k, j, i] == pytest.approx(0.4 * a, rel=1e-15)  # the centre: min(a,b,c)
    assert phi[k, j + 4, i] < 0.0 < phi[k, j + 2, i]
    # world x is the body y axis (semi-axis 0.5 a = 0.375): one cell (0.25)
    # is inside, two (0.5) outside
    assert phi[k, j, i + 2] < 0.0 < phi[k, j, i + 1]
    # no shape: -1 everywhere inside, 0 on the halo
    none = oracle_shapes(g, np.zeros((0, 10)))
    assert np.all(none[1:-1, 1:-1, 1:-1] == -1.0) and np.all(none[0] == 0.0)
