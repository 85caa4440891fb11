"""GPU parity: every entry point of the gfx950 library against the oracle
on the same seeded inputs.

Bit-exact where the library performs the reference's own arithmetic
(stencils, halos, combines, adjoint operators); 1e-12 normwise relative
error (max|gpu - oracle| / max|oracle| over the field) where the algorithm
differs in rounding order: the partition/PCR line solver against the
reference's Thomas and UL eliminations, the Stockham FFT against the
oracle's recursive FFT, GPU sin/cos/sqrt against libm.
"""
import numpy as np
import pytest

from oracle_bind import OracleRun, P, oracle, oracle_levelset
from vtkmakecompare_amd import grid as G

pytestmark = pytest.mark.gpu

TOL = 1e-12
SIZES = [(8, 9, 8), (12, 7, 24), (32, 17, 48)]


@pytest.fixture(scope="module")
def api():
    from vtkmakecompare_amd import api as A
    A.init()
    return A


def dev(a):
    from vtkmakecompare_amd.device import DeviceArray
    return DeviceArray.from_host(np.ascontiguousarray(a))


def nrel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


def metrics(g):
    return [dev(getattr(g, n)) for n in ("dx", "dy", "dz", "dx12", "dy12", "dz12")]


def rfield(rng, g, lo=-1.0, hi=1.0):
    return rng.uniform(lo, hi, size=g.shape)


@pytest.mark.parametrize("n", SIZES)
def test_set_bc(api, n):
    g = G.channel_grid(*n)
    rng = np.random.default_rng(1)
    f = [rfield(rng, g) for _ in range(5)]
    d = [dev(a) for a in f]
    api.set_bc(d[0], d[1], d[2], d[3], g.nx, g.ny, g.nz, d[4])
    oracle().o_set_bc(P(f[0]), P(f[1]), P(f[2]), P(f[3]), g.nx, g.ny, g.nz, P(f[4]))
    for a, b in zip(d, f):
        np.testing.assert_array_equal(a.download(), b)


@pytest.mark.parametrize("n", SIZES)
def test_stencils_bit_exact(api, n):
    g = G.channel_grid(*n)
    rng = np.random.default_rng(2)
    u, v, w, t = (rfield(rng, g) for _ in range(4))
    rc, th = rfield(rng, g, 0.5, 2.0), rfield(rng, g, 0.5, 2.0)
    dt, re, sc = 1e-3, 2800.0, 0.71
    M = metrics(g)
    du, dv, dw, dtt, drc, dth = (dev(a) for a in (u, v, w, t, rc, th))
    outs = [dev(np.zeros(g.shape)) for _ in range(8)]
    ref = [np.zeros(g.shape) for _ in range(8)]
    api.set_convect(du, outs[0], dv, outs[1], dw, outs[2], dt, g.nx, g.ny, g.nz, *M)
    api.set_convection_for_scalar(du, dv, dw, dt, g.nx, g.ny, g.nz, *M, dtt, outs[3], drc)
    api.set_conduction(du, outs[4], dv, outs[5], dw, outs[6], re, dt, g.nx, g.ny, g.nz, *M)
    api.set_diffusion_for_scalar(re, dt, g.nx, g.ny, g.nz, *M, dtt, outs[7], sc, dth, drc)
    O = oracle()
    ga = [P(getattr(g, n)) for n in ("dx", "dy", "dz", "dx12", "dy12", "dz12")]
    O.o_set_convect(P(u), P(ref[0]), P(v), P(ref[1]), P(w), P(ref[2]), dt, g.nx, g.ny, g.nz, *ga)
    O.o_set_convection_for_scalar(P(u), P(v), P(w), dt, g.nx, g.ny, g.nz, *ga, P(t), P(ref[3]),
                                  P(rc))
    O.o_set_conduction(P(u), P(ref[4]), P(v), P(ref[5]), P(w), P(ref[6]), re, dt, g.nx, g.ny,
                       g.nz, *ga)
    O.o_set_diffusion_for_scalar(re, dt, g.nx, g.ny, g.nz, *ga, P(t), P(ref[7]), sc, P(th), P(rc))
    for a, b in zip(outs, ref):
        np.testing.assert_array_equal(interior(a.download()), interior(b))


@pytest.mark.parametrize("l", [1, 2])
def test_immersed_explicit_bit_exact(api, l):
    g = G.channel_grid(12, 9, 8)
    rng = np.random.default_rng(10 + l)
    cv = np.zeros((4, g.nz, g.ny, g.nx))
    cv[:] = rng.uniform(0, 1, cv.shape)
    uta, vta, wta = (rfield(rng, g) for _ in range(3))
    eta, dt = 1e4, 1e-3
    dcv = dev(cv)
    M = metrics(g)
    dl = [dev(np.zeros(g.shape)) for _ in range(4)]
    api.set_immersed_boundary(eta, g.nx, g.ny, g.nz, dt, dl[0], dl[1], dl[2], M[0], M[1], M[2],
                              dl[3], dcv, dev(uta), dev(vta), dev(wta), l)
    rl = [np.zeros(g.shape) for _ in range(4)]
    cons = np.zeros(1)
    cvo = cv.copy()
    oracle().o_set_immersed_boundary(eta, g.nx, g.ny, g.nz, dt, P(rl[0]), P(rl[1]), P(rl[2]),
                                     P(g.dx), P(g.dy), P(g.dz), P(rl[3]), P(cvo), P(uta),
                                     P(vta), P(wta), l, P(cons))
    np.testing.assert_array_equal(dcv.download(), cvo)
    for a, b in zip(dl, rl):
        np.testing.assert_array_equal(interior(a.download()), interior(b))
    ins = [rfield(rng, g) for _ in range(15)]
    d = [dev(a) for a in ins]
    api.set_explicit(*d, g.nx, g.ny, g.nz, l)
    oracle().o_set_explicit(*[P(a) for a in ins], g.nx, g.ny, g.nz, l)
    for a, b in zip(d, ins):
        np.testing.assert_array_equal(a.download(), b)
    tin = [rfield(rng, g) for _ in range(5)]
    dt_ = [dev(a) for a in tin]
    api.set_explicit_for_tempreture(*dt_, g.nx, g.ny, g.nz, l)
    oracle().o_set_explicit_for_tempreture(*[P(a) for a in tin], g.nx, g.ny, g.nz, l)
    for a, b in zip(dt_, tin):
        np.testing.assert_array_equal(a.download(), b)
    # main-loop T divide
    ut, t = rfield(rng, g), rfield(rng, g)
    dut = dev(ut)
    api.t_divide(dut, dev(t), dcv, eta, dt, g.nx, g.ny, g.nz)
    oracle().o_t_divide(P(ut), P(t), P(cvo), eta, dt, g.nx, g.ny, g.nz)
    np.testing.assert_array_equal(dut.download(), ut)


def sparse_cellvalue(rng, g):
    """cellvalue zero except a blob in one corner: most lines are fluid-only
    (their matrices equal the reference line's), a few are not."""
    cv = np.zeros((4, g.nz, g.ny, g.nx))
    kz, jy, ix = max(1, g.nz // 3), max(1, g.ny // 3), max(1, g.nx // 5)
    cv[:, :kz, :jy, :ix] = rng.uniform(0, 1, (4, kz, jy, ix))
    return cv


@pytest.mark.parametrize("sparse", [False, True])
@pytest.mark.parametrize("n", SIZES + [(64, 65, 64)])
def test_solve_u_temp(api, n, sparse):
    g = G.channel_grid(*n)
    rng = np.random.default_rng(3)
    cv = sparse_cellvalue(rng, g) if sparse else rng.uniform(0, 1, (4, g.nz, g.ny, g.nx))
    a = [rfield(rng, g) for _ in range(3)]
    old = [rfield(rng, g) for _ in range(3)]
    re, dt, eta = 2800.0, 1e-3, 1e4
    da = [dev(x) for x in a]
    M = metrics(g)
    api.solve_u_temp(da[0], da[1], da[2], g.nx, g.ny, g.nz, *M, re, dt, eta, dev(cv),
                     *[dev(x) for x in old])
    ga = [P(getattr(g, n)) for n in ("dx", "dy", "dz", "dx12", "dy12", "dz12")]
    oracle().o_solve_u_temp(P(a[0]), P(a[1]), P(a[2]), g.nx, g.ny, g.nz, *ga, re, dt, eta,
                            P(cv), P(old[0]), P(old[1]), P(old[2]))
    for x, y in zip(da, a):
        got = x.download()
        # interior and every halo cell the sweeps write
        assert nrel(interior(got), interior(y)) < TOL
        assert nrel(got[1:-1, :, 1:-1], y[1:-1, :, 1:-1]) < TOL


@pytest.mark.parametrize("sparse", [False, True])
@pytest.mark.parametrize("n", SIZES + [(64, 65, 64)])
def test_solve_t(api, n, sparse):
    g = G.channel_grid(*n)
    rng = np.random.default_rng(4)
    ut = rfield(rng, g)
    rc, th = np.ones(g.shape), np.ones(g.shape)
    if sparse:
        cv = sparse_cellvalue(rng, g)
    else:
        cv = rng.uniform(0, 1, (4, g.nz, g.ny, g.nx))
        rc[1:-1, 1:-1, 1:-1] = rng.uniform(0.5, 2, (g.nz, g.ny, g.nx))
        th[1:-1, 1:-1, 1:-1] = rng.uniform(0.5, 2, (g.nz, g.ny, g.nx))
    re, sc, dt, eta = 2800.0, 0.71, 1e-3, 1e4
    d = dev(ut)
    M = metrics(g)
    api.solve_t(d, g.nx, g.ny, g.nz, *M, re, sc, dt, dev(rc), dev(th), dev(cv), eta)
    ga = [P(getattr(g, n)) for n in ("dx", "dy", "dz", "dx12", "dy12", "dz12")]
    oracle().o_solve_t(P(ut), g.nx, g.ny, g.nz, *ga, re, sc, dt, P(rc), P(th), P(cv), eta)
    assert nrel(interior(d.download()), interior(ut)) < TOL


@pytest.mark.parametrize("n", [(8, 5, 6), (12, 3, 24), (16, 4, 8), (384, 3, 384)])
def test_p2s_s2p(api, n):
    nx, ny, nz = n
    g = G.channel_grid(nx, ny, nz)
    rng = np.random.default_rng(nx)
    fp = rfield(rng, g)
    nh = nx // 2
    dfp = dev(fp)
    dft = dev(np.zeros((nz + 2, ny + 2, nh + 1, 2)))
    api.p2s(dfp, dft, nx, ny, nz)
    ref = np.zeros((nz + 2, ny + 2, nh + 1, 2))
    oracle().o_p2s(P(fp), P(ref), nx, ny, nz)
    got = dft.download()
    assert nrel(got[0:nz, 1:ny + 1], ref[0:nz, 1:ny + 1]) < TOL
    back = dev(np.zeros(g.shape))
    api.s2p(dft, back, nx, ny, nz)
    rb = np.zeros(g.shape)
    oracle().o_s2p(P(ref), P(rb), nx, ny, nz)
    gb = back.download()
    assert nrel(gb[0:nz + 1, 1:ny + 1, 0:nx + 1], rb[0:nz + 1, 1:ny + 1, 0:nx + 1]) < TOL
    # S2P leaves its input untouched
    np.testing.assert_array_equal(dft.download(), got)


@pytest.mark.parametrize("n", [(8, 9, 6), (32, 33, 16), (64, 129, 32)])
def test_spn_sp0b(api, n):
    nx, ny, nz = n
    g = G.channel_grid(nx, ny, nz)
    rng = np.random.default_rng(5)
    nh = nx // 2
    kxz = np.zeros((nz + 1, nh + 1))
    ac = np.zeros((3, ny))
    oracle().o_mkp_coefficients(nx, ny, nz, P(g.dx), P(g.dy), P(g.dz), P(g.dy12), g.xl, g.zl,
                                P(kxz), P(ac))
    phit = rng.normal(size=(nz + 2, ny + 2, nh + 1, 2))
    d = dev(phit)
    api.spn(nx, ny, nz, dev(ac), dev(kxz), d)
    api.sp0b(nx, ny, nz, dev(ac), d)
    oracle().o_spn(nx, ny, nz, P(ac), P(kxz), P(phit))
    oracle().o_sp0b(nx, ny, nz, P(ac), P(phit))
    got = d.download()
    assert nrel(got[0:nz, 1:ny + 1], phit[0:nz, 1:ny + 1]) < TOL
    np.testing.assert_allclose(got[0, :, 0], phit[0, :, 0], rtol=0,
                               atol=TOL * np.max(np.abs(phit[0, :, 0])))


def _mkp_inputs(g, rng):
    ut = [rfield(rng, g) for _ in range(3)]
    p = rfield(rng, g)
    t = np.zeros(g.shape)
    oracle().o_set_bc(P(ut[0]), P(ut[1]), P(ut[2]), P(p), g.nx, g.ny, g.nz, P(t))
    return ut, p


MKP_OUT = ("u_ta", "v_ta", "w_ta", "u_ep2", "v_ep2", "w_ep2", "divu", "u", "v", "w")


@pytest.mark.parametrize("n", [(8, 9, 6), (12, 17, 24), (64, 65, 32)])
def test_mkp(api, n):
    g = G.channel_grid(*n)
    rng = np.random.default_rng(6)
    ut, p = _mkp_inputs(g, rng)
    outs = {k: rfield(rng, g) for k in MKP_OUT}
    d = {k: dev(v) for k, v in outs.items()}
    dp = dev(p)
    dt = 1e-3
    M = metrics(g)
    api.mkp(dp, *[dev(x) for x in ut], dt, *M, 2800.0, g.nx, g.ny, g.nz, 1,
            *[d[k] for k in MKP_OUT], g.xl, g.zl)
    ga = [P(getattr(g, n)) for n in ("dx", "dy", "dz", "dx12", "dy12", "dz12")]
    oracle().o_mkp(P(p), P(ut[0]), P(ut[1]), P(ut[2]), dt, *ga, 2800.0, g.nx, g.ny, g.nz,
                   *[P(outs[k]) for k in MKP_OUT], g.xl, g.zl)
    assert nrel(interior(dp.download()), interior(p)) < TOL
    for k in ("u", "v", "w", "u_ta", "v_ta", "w_ta", "u_ep2", "v_ep2", "w_ep2"):
        assert nrel(interior(d[k].download()), interior(outs[k])) < TOL, k
    # divergence: absolute, both at round-off
    scale = max(np.max(np.abs(interior(outs[k]))) for k in ("u", "v", "w")) / g.dx[1]
    assert np.max(np.abs(interior(d["divu"].download()))) < 1e-11 * scale


def test_levelset(api):
    g = G.channel_grid(16, 17, 16)
    phi0 = G.sphere_levelset(g, [(np.pi, 0.1, 0.5 * np.pi, 0.7)])
    phi_o, cv_o = oracle_levelset(g, phi0, 2001)
    dphi, dcv = dev(np.zeros(g.shape)), dev(np.zeros((4, g.nz, g.ny, g.nx)))
    M = [dev(getattr(g, n)) for n in ("dx", "dx12", "dy12", "dz12")]
    api.set_levelset(g.nx, g.ny, g.nz, *M, dev(phi0), dphi, dev(g.x), dev(g.y), dev(g.z), dcv)
    assert nrel(dphi.download(), phi_o) < TOL
    assert nrel(dcv.download()[0], cv_o[0]) < TOL


def test_adjoint_kernels_bit_exact(api):
    g = G.channel_grid(12, 9, 8)
    rng = np.random.default_rng(9)
    u, v, w, uf, vf, wf, t, tf = (rfield(rng, g) for _ in range(8))
    rc = rfield(rng, g, 0.5, 2.0)
    re, dt = 2800.0, 1e-3
    M = metrics(g)
    D = {k: dev(x) for k, x in dict(u=u, v=v, w=w, uf=uf, vf=vf, wf=wf, t=t, tf=tf, rc=rc).items()}
    out = [dev(np.zeros(g.shape)) for _ in range(7)]
    api.set_adj_convect(D["u"], out[0], D["v"], out[1], D["w"], out[2], re, dt, g.nx, g.ny,
                        g.nz, *M, D["uf"], D["vf"], D["wf"])
    api.set_adj_convection_for_scalar(D["u"], D["v"], D["w"], re, dt, g.nx, g.ny, g.nz, *M,
                                      D["t"], out[3], D["rc"], D["uf"], D["vf"], D["wf"])
    api.set_adj_ns_scalar(dt, g.nx, g.ny, g.nz, M[3], M[4], M[5], D["t"], out[4], out[5],
                          out[6], D["uf"], D["vf"], D["wf"], D["tf"])
    ref = [np.zeros(g.shape) for _ in range(7)]
    ga = [P(getattr(g, n)) for n in ("dx", "dy", "dz", "dx12", "dy12", "dz12")]
    O = oracle()
    O.o_set_adj_convect(P(u), P(ref[0]), P(v), P(ref[1]), P(w), P(ref[2]), re, dt, g.nx, g.ny,
                        g.nz, *ga, P(uf), P(vf), P(wf))
    O.o_set_adj_convection_for_scalar(P(u), P(v), P(w), re, dt, g.nx, g.ny, g.nz, *ga, P(t),
                                      P(ref[3]), P(rc), P(uf), P(vf), P(wf))
    O.o_set_adj_ns_scalar(dt, g.nx, g.ny, g.nz, ga[3], ga[4], ga[5], P(t), P(ref[4]), P(ref[5]),
                          P(ref[6]), P(uf), P(vf), P(wf), P(tf))
    for a, b in zip(out, ref):
        np.testing.assert_array_equal(interior(a.download()), interior(b))


def _small_case(n, seed=0x5EED00AA):
    cfg = G.small_config(*n, seed=seed)
    g = cfg.grid()
    st = G.initial_state(cfg, g)
    _, cv = oracle_levelset(g, st["phi0"], 200)
    return cfg, g, st, cv


@pytest.mark.parametrize("n", [(16, 17, 16), (24, 33, 12)])
def test_time_step_matches_oracle(api, n):
    cfg, g, st, cv = _small_case(n)
    run = api.ChannelRun(cfg, g, st, cellvalue=cv)
    st_o = dict(st)
    st_o["cellvalue"] = cv.copy()
    ora = OracleRun(cfg, g, st_o)
    nsteps = 5
    run.step(nsteps)
    ora.step(nsteps)
    for k in ("u", "v", "w", "p", "t", "u_ta", "v_ta", "w_ta", "u_pre", "v_pre", "w_pre",
              "t_pre", "u_ep2", "v_ep2", "w_ep2"):
        assert nrel(interior(run.get(k)), interior(ora[k])) < TOL, k
    for k in ("u_temp", "v_temp", "w_temp"):
        assert nrel(run.get(k), ora[k]) < TOL, k
    div = np.max(np.abs(interior(run.get("divu"))))
    umax = np.max(np.abs(interior(ora["u"])))
    assert div < 1e-10 * umax / g.dy[1]
