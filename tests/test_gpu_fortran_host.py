"""The Fortran drop-in host (vtkmakecompare_amd/host/vk_driver): program
main's time loop calling the reference routine names, on the GPU, against
the oracle.  Mode 0 is the reference call sequence through the wrappers,
mode 1 the fused step; both must match the oracle to 1e-12.  Mode 2 is the
mixed-precision step (velocity 1e-6, p and T 1e-5 against the oracle);
mode 3 the objective and adjoint gradient through the Fortran bindings,
equal to the same computation through the Python host."""
import os
import subprocess

import numpy as np
import pytest

from oracle_bind import OracleRun, oracle_levelset
from vtkmakecompare_amd import grid as G

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "vtkmakecompare_amd", "host", "vk_driver")
OUT = ("u", "v", "w", "p", "t", "u_ta", "v_ta", "w_ta", "divu", "u_temp", "v_temp", "w_temp")


def write_input(path, cfg, g, st, cv, nsteps, mode, every=0, levelset=0, vtk=0):
    with open(path, "wb") as f:
        np.array([g.nx, g.ny, g.nz, nsteps, mode, levelset, every, vtk],
                 dtype=np.int32).tofile(f)
        np.array([cfg.dt, cfg.re, cfg.sc, cfg.eta, g.xl, g.zl, g.yl, 0],
                 dtype=np.float64).tofile(f)
        for n in ("dx", "dy", "dz", "dx12", "dy12", "dz12", "x", "y", "z"):
            np.ascontiguousarray(getattr(g, n), dtype=np.float64).tofile(f)
        for n in ("u", "v", "w", "t", "phi0"):
            np.ascontiguousarray(st[n]).tofile(f)
        np.ascontiguousarray(cv).tofile(f)


def read_output(path, g):
    raw = np.fromfile(path, dtype=np.float64)
    nf = int(np.prod(g.shape))
    out = {}
    for q, n in enumerate(OUT):
        out[n] = raw[q * nf:(q + 1) * nf].reshape(g.shape)
    out["cellvalue"] = raw[len(OUT) * nf:].reshape(4, g.nz, g.ny, g.nx)
    return out


def nrel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran host not built")
@pytest.mark.parametrize("mode", [0, 1])
def test_fortran_driver_matches_oracle(tmp_path, mode):
    cfg = G.small_config(16, 17, 16, seed=0x5EED00BB)
    g = cfg.grid()
    st = G.initial_state(cfg, g)
    _, cv = oracle_levelset(g, st["phi0"], 100)
    nsteps = 4
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(fin, cfg, g, st, cv, nsteps, mode)
    print(f"vk_driver mode {mode} 16x17x16 {nsteps} steps", flush=True)
    # in tmp_path: the CHECK_CONVERGENCE wrapper appends to timely_convergence.dat
    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr + r.stdout
    got = read_output(fout, g)
    st_o = dict(st)
    st_o["cellvalue"] = cv.copy()
    ora = OracleRun(cfg, g, st_o)
    ora.step(nsteps)
    for n in ("u", "v", "w", "p", "t", "u_ta", "v_ta", "w_ta"):
        assert nrel(got[n][1:-1, 1:-1, 1:-1], ora[n][1:-1, 1:-1, 1:-1]) < 1e-12, n
    for n in ("u_temp", "v_temp", "w_temp"):
        assert nrel(got[n], ora[n]) < 1e-12, n
    np.testing.assert_array_equal(got["cellvalue"], ora["cellvalue"])
    check_diagnostics(r.stdout, tmp_path, cfg, g, ora, nsteps)


def check_diagnostics(stdout, tmp_path, cfg, g, ora, nsteps):
    """The reference-named diagnostics wrappers on the final state against
    the oracle's (the fields agree to 1e-12, the sums' order differs)."""
    import ctypes
    from oracle_bind import P, oracle
    lines = stdout.splitlines()
    cour = float(next(x for x in lines if "courant max =" in x).split()[-1])
    c = ctypes.c_double(0.0)
    oracle().o_check_courant(cfg.dt, P(g.dx), P(g.dy), P(g.dz), g.nx, g.ny, g.nz,
                             P(ora["u"]), P(ora["v"]), P(ora["w"]), ctypes.byref(c))
    assert cour == pytest.approx(c.value, rel=1e-10)
    mon = [float(x) for x in next(x for x in lines if "vk_driver: monitor" in x).split()[2:]]
    o = {k: ctypes.c_double(0.0) for k in ("dif", "ta", "td", "vd", "ud", "vdd", "wd", "ub",
                                          "vb", "wb", "tb")}
    loc = (ctypes.c_int * 3)()
    oracle().o_monitor(g.nx, g.ny, g.nz, 1, P(ora["t"]), P(g.dx), P(g.dy), P(g.dz),
                       *[ctypes.byref(o[k]) for k in ("dif", "ta", "td")], P(ora["t_pre"]),
                       P(ora["u"]), P(ora["v"]), P(ora["w"]),
                       *[ctypes.byref(o[k]) for k in ("vd", "ud", "vdd", "wd", "ub", "vb", "wb",
                                                     "tb")], cfg.dt, loc)
    assert mon[0] == pytest.approx(o["ta"].value, rel=1e-10)
    assert mon[1] == pytest.approx(o["td"].value, rel=1e-9, abs=1e-11)
    assert mon[2] == pytest.approx(o["ub"].value, rel=1e-10)
    rec = (tmp_path / "timely_convergence.dat").read_text().split()
    e = ctypes.c_double(0.0)
    oracle().o_check_convergence(P(g.y), g.yl, P(ora["u"]), cfg.re, g.nx, g.ny, g.nz,
                                 ctypes.byref(e))
    assert float(rec[0]) == pytest.approx(nsteps * cfg.dt, rel=1e-12)
    assert float(rec[1]) == pytest.approx(e.value, rel=1e-10)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran host not built")
def test_fortran_driver_mixed_step(tmp_path):
    cfg = G.small_config(24, 17, 24, seed=0x5EED00BC)
    g = cfg.grid()
    st = G.initial_state(cfg, g)
    _, cv = oracle_levelset(g, st["phi0"], 100)
    nsteps = 5
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(fin, cfg, g, st, cv, nsteps, 2)
    print(f"vk_driver mode 2 24x17x24 {nsteps} steps", flush=True)
    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    got = read_output(fout, g)
    st_o = dict(st)
    st_o["cellvalue"] = cv.copy()
    ora = OracleRun(cfg, g, st_o)
    ora.step(nsteps)
    for n in ("u", "v", "w"):
        assert nrel(got[n][1:-1, 1:-1, 1:-1], ora[n][1:-1, 1:-1, 1:-1]) < 1e-6, n
    for n in ("p", "t", "u_ta", "v_ta", "w_ta"):
        assert nrel(got[n][1:-1, 1:-1, 1:-1], ora[n][1:-1, 1:-1, 1:-1]) < 1e-5, n
    assert float(np.max(np.abs(got["divu"][1:-1, 1:-1, 1:-1]))) < 1e-10


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran host not built")
def test_fortran_driver_adjoint(tmp_path):
    from vtkmakecompare_amd import api
    api.init()
    cfg = G.small_config(16, 17, 16, seed=0x5EED00BD)
    g = cfg.grid()
    st = G.initial_state(cfg, g)
    _, cv = oracle_levelset(g, st["phi0"], 100)
    nsteps, every = 4, 3
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(fin, cfg, g, st, cv, nsteps, 3, every)
    print(f"vk_driver mode 3 16x17x16 {nsteps} steps", flush=True)
    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    raw = np.fromfile(fout, dtype=np.float64)
    j_f = raw[0]
    grad_f = raw[1:].reshape(g.nz, g.ny, g.nx)
    prob = api.AdjointProblem(cfg, g, st, cv, nsteps)
    j_p, grad_p = prob.gradient(checkpoint_every=every)
    assert abs(j_f - j_p) <= 1e-14 * abs(j_p)
    assert nrel(grad_f, grad_p) <= 1e-14


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran host not built")
def test_fortran_component_routines(tmp_path):
    """Mode 4: the projection's and the adjoint's component routines through
    the Fortran wrappers with the reference's names and argument lists
    (P2S, S2P, SPN, SP0B, PHIBC, CTRIDIAG, set_adj_convect,
    set_adj_convection_for_scalar, set_adj_NS_scalar), against the oracle
    on the same inputs: the FFT-based results to 1e-12 normwise, the rest
    bit for bit."""
    from oracle_bind import P, oracle
    cfg = G.small_config(16, 17, 12, seed=0x5EED00BE)
    g = cfg.grid()
    nx, ny, nz = g.nx, g.ny, g.nz
    rng = np.random.default_rng(41)
    st = {k: rng.normal(size=g.shape) for k in ("u", "v", "w", "t", "phi0")}
    cv = np.zeros((4, nz, ny, nx))
    o = oracle()
    kxz = np.zeros((nz + 1, nx // 2 + 1))
    ac = np.zeros((3, ny))
    o.o_mkp_coefficients(nx, ny, nz, P(g.dx), P(g.dy), P(g.dz), P(g.dy12), g.xl, g.zl, P(kxz),
                         P(ac))
    atri = np.empty((3, ny))
    atri[0], atri[2] = rng.uniform(-1, 0, size=ny), rng.uniform(-1, 0, size=ny)
    atri[1] = -3.0 - rng.uniform(0, 1, size=ny)
    btri = rng.normal(size=(ny, 2))
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(fin, cfg, g, st, cv, 0, 4)
    with open(fin, "ab") as f:
        for a in (ac, kxz, atri, btri):
            np.ascontiguousarray(a).tofile(f)
    print("vk_driver mode 4 16x17x12", flush=True)
    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    raw = np.fromfile(fout, dtype=np.float64)
    nf = int(np.prod(g.shape))
    ns = 2 * (nx // 2 + 1) * (ny + 2) * (nz + 2)
    sizes = [("fpt", ns), ("phi", nf), ("atri", 3 * ny), ("btri", 2 * ny), ("u_v", nf),
             ("v_v", nf), ("w_v", nf), ("u_vt", nf), ("u_s", nf), ("v_s", nf), ("w_s", nf)]
    got, q = {}, 0
    for name, n in sizes:
        got[name] = raw[q:q + n]
        q += n
    assert q == raw.size
    # the oracle on the same inputs
    fpt = np.zeros(ns)
    o.o_p2s(P(st["u"]), P(fpt), nx, ny, nz)
    assert nrel(got["fpt"], fpt) < 1e-12
    fpt2 = np.zeros(ns)
    o.o_p2s(P(st["v"]), P(fpt2), nx, ny, nz)
    o.o_spn(nx, ny, nz, P(ac), P(kxz), P(fpt2))
    o.o_sp0b(nx, ny, nz, P(ac), P(fpt2))
    phi = np.zeros(g.shape)
    o.o_s2p(P(fpt2), P(phi), nx, ny, nz)
    o.o_phibc(nx, ny, nz, P(phi))
    gphi = got["phi"].reshape(g.shape)
    assert nrel(gphi[1:-1, 1:-1, 1:-1], phi[1:-1, 1:-1, 1:-1]) < 1e-12
    a, b = atri.copy(), np.ascontiguousarray(btri).copy()
    o.o_ctridiag(P(a), P(b), ny)
    np.testing.assert_array_equal(got["atri"], a.ravel())
    np.testing.assert_array_equal(got["btri"], b.ravel())
    u, v, w, t, tf = (st[k] for k in ("u", "v", "w", "t", "phi0"))
    rc = np.ones(g.shape)
    met = [P(getattr(g, k)) for k in ("dx", "dy", "dz", "dx12", "dy12", "dz12")]
    uv, vv, wv, uvt = (np.zeros(g.shape) for _ in range(4))
    o.o_set_adj_convect(P(u), P(uv), P(v), P(vv), P(w), P(wv), cfg.re, cfg.dt, nx, ny, nz, *met,
                        P(t), P(w), P(v))
    o.o_set_adj_convection_for_scalar(P(u), P(v), P(w), cfg.re, cfg.dt, nx, ny, nz, *met, P(t),
                                      P(uvt), P(rc), P(t), P(w), P(v))
    us, vs, ws = (np.zeros(g.shape) for _ in range(3))
    o.o_set_adj_ns_scalar(cfg.dt, nx, ny, nz, met[3], met[4], met[5], P(t), P(us), P(vs), P(ws),
                          P(t), P(w), P(v), P(tf))
    inner = (slice(1, -1),) * 3
    for name, want in (("u_v", uv), ("v_v", vv), ("w_v", wv), ("u_vt", uvt), ("u_s", us),
                       ("v_s", vs), ("w_s", ws)):
        np.testing.assert_array_equal(got[name].reshape(g.shape)[inner], want[inner],
                                      err_msg=name)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran host not built")
def test_fortran_set_levelset_writes_the_vtk_files(tmp_path):
    """With the VTK flag, the set_levelset wrapper writes levelset.vtk and
    cellvalue1.vtk and the driver cellvalue.vtk, as the reference's
    set_levelset and program main do: cellvalue(...,1) byte for byte as the
    library's writer gives it from the run's cellvalue (whose format is
    pinned against the reference's own files on the CPU)."""
    from vtkmakecompare_amd import api
    cfg = G.small_config(16, 17, 12, seed=0x5EED00BF)
    g = cfg.grid()
    st = G.initial_state(cfg, g)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(fin, cfg, g, st, np.zeros((4, g.nz, g.ny, g.nx)), 1, 1, levelset=1, vtk=1)
    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120,
                       cwd=tmp_path)
    assert r.returncode == 0, r.stderr + r.stdout
    got = read_output(fout, g)
    (tmp_path / "py").mkdir()
    cwd = os.getcwd()
    os.chdir(tmp_path / "py")
    try:
        api.vtk_cellvalue(g, np.ascontiguousarray(got["cellvalue"]))
    finally:
        os.chdir(cwd)
    want = (tmp_path / "py" / "cellvalue.vtk").read_bytes()
    assert (tmp_path / "cellvalue1.vtk").read_bytes() == want
    assert (tmp_path / "cellvalue.vtk").read_bytes() == want
    lev = (tmp_path / "levelset.vtk").read_text().splitlines()
    n = g.nx * g.ny * g.nz
    assert lev[4] == f"DIMENSIONS  {g.nx:3d} {g.ny:3d} {g.nz:3d}"
    assert lev[6 + n + 1] == "SCALARS phi float64" and len(lev) == 6 + n + 3 + n
