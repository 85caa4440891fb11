"""Python host mirror of the reference's routine interface.

Each function has the name and the argument list of the Fortran routine it
replaces (see the citations in include/vtkgpu.h); arrays are
:class:`~vtkmakecompare_amd.device.DeviceArray` objects in the reference
layout, scalars are Python numbers.  Every call runs the gfx950 library;
a non-zero status raises :class:`VKError`.

:class:`ChannelRun` holds a whole configuration resident in HBM and
advances it with the fused time step (the loop body of program main).
"""
from __future__ import annotations

import numpy as np

from . import grid as G
from ._lib import D, I, VKError, check, lib  # noqa: F401
from .device import DeviceArray, init, synchronize  # noqa: F401

NULL = None


def set_bc(u, v, w, p, nx, ny, nz, t):
    check(lib().vk_set_bc(u, v, w, p, I(nx), I(ny), I(nz), t), "SET_BC")


def set_convect(u, u_v, v, v_v, w, w_v, dt, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12):
    check(lib().vk_set_convect(u, u_v, v, v_v, w, w_v, D(dt), I(nx), I(ny), I(nz),
                               dx, dy, dz, dx12, dy12, dz12), "SET_CONVECT")


def set_convection_for_scalar(u, v, w, dt, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12,
                              t, u_vt, rhocp):
    check(lib().vk_set_convection_for_scalar(u, v, w, D(dt), I(nx), I(ny), I(nz), dx, dy, dz,
                                             dx12, dy12, dz12, t, u_vt, rhocp),
          "SET_CONVECTION_FOR_SCALAR")


def set_conduction(u, u_d, v, v_d, w, w_d, re, dt, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12):
    check(lib().vk_set_conduction(u, u_d, v, v_d, w, w_d, D(re), D(dt), I(nx), I(ny), I(nz),
                                  dx, dy, dz, dx12, dy12, dz12), "SET_CONDUCTION")


def set_diffusion_for_scalar(re, dt, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12, t, u_dt, sc,
                             thc, rhocp):
    check(lib().vk_set_diffusion_for_scalar(D(re), D(dt), I(nx), I(ny), I(nz), dx, dy, dz,
                                            dx12, dy12, dz12, t, u_dt, D(sc), thc, rhocp),
          "SET_DIFFUSION_FOR_SCALAR")


def set_immersed_boundary(eta, nx, ny, nz, dt, u_l, v_l, w_l, dx, dy, dz, t_l, cellvalue,
                          u_ta, v_ta, w_ta, l):
    import ctypes
    cons = ctypes.c_double(0.0)
    check(lib().vk_set_immersed_boundary(D(eta), I(nx), I(ny), I(nz), D(dt), u_l, v_l, w_l,
                                         dx, dy, dz, t_l, cellvalue, u_ta, v_ta, w_ta, I(l),
                                         ctypes.byref(cons)), "SET_IMMERSED_BOUNDARY")
    return cons.value


def set_explicit(u_v, u_pre, u_d, u_a, v_v, v_pre, v_d, v_a, w_v, w_pre, w_d, w_a,
                 u_l, v_l, w_l, nx, ny, nz, l):
    check(lib().vk_set_explicit(u_v, u_pre, u_d, u_a, v_v, v_pre, v_d, v_a, w_v, w_pre, w_d,
                                w_a, u_l, v_l, w_l, I(nx), I(ny), I(nz), I(l)), "SET_EXPLICIT")


def set_explicit_for_tempreture(u_vt, u_dt, u_t, t_pre, t_l, nx, ny, nz, l):
    check(lib().vk_set_explicit_for_tempreture(u_vt, u_dt, u_t, t_pre, t_l, I(nx), I(ny),
                                               I(nz), I(l)), "set_explicit_for_tempreture")


def t_divide(u_t, t, cellvalue, eta, dt, nx, ny, nz):
    check(lib().vk_t_divide(u_t, t, cellvalue, D(eta), D(dt), I(nx), I(ny), I(nz)),
          "T divide")


def solve_u_temp(u_a, v_a, w_a, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12, re, dt, eta,
                 cellvalue, u, v, w):
    check(lib().vk_solve_u_temp(u_a, v_a, w_a, I(nx), I(ny), I(nz), dx, dy, dz, dx12, dy12,
                                dz12, D(re), D(dt), D(eta), cellvalue, u, v, w),
          "solve_u_temp")


def solve_t(u_t, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12, re, sc, dt, rhocp, thc,
            cellvalue, eta):
    check(lib().vk_solve_t(u_t, I(nx), I(ny), I(nz), dx, dy, dz, dx12, dy12, dz12, D(re),
                           D(sc), D(dt), rhocp, thc, cellvalue, D(eta)), "solve_t")


# ------------------------------------------- batched line-level solvers
# Line arrays are (n+2, nsys) C-order arrays on the device (Fortran
# X(nsys, 0:n+1), the systems fastest); per-line scalars are (nsys,)
# device arrays.

def ctridiag_batched(a, b, nn, nsys):
    """CTRIDIAG (tridiag.f:43): a (3, nn, nsys), b (nn, nsys) complex as
    (nn, nsys, 2) doubles, in place."""
    check(lib().vk_ctridiag_batched(a, b, I(nn), I(nsys)), "ctridiag_batched")


def tdma_batched(a, b, c, r, u, n, nsys):
    """TDMA (periodic_implicit.f:47)."""
    check(lib().vk_tdma_batched(a, b, c, r, u, I(n), I(nsys)), "tdma_batched")


def cyclic_batched(a, b, c, alpha, beta, r, x, n, nsys):
    """cyclic (periodic_implicit.f:67)."""
    check(lib().vk_cyclic_batched(a, b, c, alpha, beta, r, x, I(n), I(nsys)), "cyclic_batched")


def implicit_solver_batched(sux, nx, dx, dx12, re, dt, ft, fb, eta, cux, nsys):
    """implicit_solver (periodic_implicit.f:1)."""
    check(lib().vk_implicit_solver_batched(sux, I(nx), dx, dx12, D(re), D(dt), ft, fb, D(eta),
                                           cux, I(nsys)), "implicit_solver_batched")


def implicit_uy_batched(s, ny, dy, dy12, re, dt, ft, fb, eta, c, nsys):
    """implicit_uy (nse_dominik.f:1505)."""
    check(lib().vk_implicit_uy_batched(s, I(ny), dy, dy12, D(re), D(dt), ft, fb, D(eta), c,
                                       I(nsys)), "implicit_uy_batched")


def implicit_vy_batched(s, ny, dy, dy12, re, dt, ft, fb, eta, c, nsys):
    """implicit_vy (nse_dominik.f:1583)."""
    check(lib().vk_implicit_vy_batched(s, I(ny), dy, dy12, D(re), D(dt), ft, fb, D(eta), c,
                                       I(nsys)), "implicit_vy_batched")


def implicit_tx_batched(stx, nx, dx, dx12, re, sc, dt, llx, rrx, ft, fb, eta, cux, nsys):
    """implicit_Tx (nse_dominik.f:1725)."""
    check(lib().vk_implicit_tx_batched(stx, I(nx), dx, dx12, D(re), D(sc), D(dt), llx, rrx, ft,
                                       fb, D(eta), cux, I(nsys)), "implicit_tx_batched")


def implicit_ty_batched(sty, ny, dy, dy12, re, sc, dt, lly, rry, ft, fb, cuy, eta, nsys):
    """implicit_Ty (nse_dominik.f:1784)."""
    check(lib().vk_implicit_ty_batched(sty, I(ny), dy, dy12, D(re), D(sc), D(dt), lly, rry, ft,
                                       fb, cuy, D(eta), I(nsys)), "implicit_ty_batched")


def linebyline(nx, ny, nz, u_temp, v_temp, w_temp, dx, dy, dz, dx12, dy12, dz12, dt, p):
    """linebyline (nse_dominik.f:2136): one line Gauss-Seidel iteration."""
    check(lib().vk_linebyline(I(nx), I(ny), I(nz), u_temp, v_temp, w_temp, dx, dy, dz, dx12,
                              dy12, dz12, D(dt), p), "linebyline")


def solve_p(p, u_temp, v_temp, w_temp, dt, dx, dy, dz, dx12, dy12, dz12, nx, ny, nz,
            max_iterations=None):
    """solve_p (nse_dominik.f:1955): returns (res, res_max, res_int, res_var,
    iterations)."""
    import ctypes
    r = [ctypes.c_double(0.0) for _ in range(4)]
    it = ctypes.c_int(0)
    mi = None if max_iterations is None else I(max_iterations)
    check(lib().vk_solve_p(p, u_temp, v_temp, w_temp, D(dt), dx, dy, dz, dx12, dy12, dz12,
                           I(nx), I(ny), I(nz), *[ctypes.byref(x) for x in r], mi,
                           ctypes.byref(it)), "solve_p")
    return tuple(x.value for x in r) + (it.value,)


def mkp(p, u_temp, v_temp, w_temp, dt, dx, dy, dz, dx12, dy12, dz12, re, nx, ny, nz,
        ompnthreads, u_ta, v_ta, w_ta, u_ep2, v_ep2, w_ep2, divu, u, v, w, xl, zl):
    check(lib().vk_mkp(p, u_temp, v_temp, w_temp, D(dt), dx, dy, dz, dx12, dy12, dz12, D(re),
                       I(nx), I(ny), I(nz), I(ompnthreads), u_ta, v_ta, w_ta, u_ep2, v_ep2,
                       w_ep2, divu, u, v, w, D(xl), D(zl)), "MKP")


def p2s(fp, fpt, nx, ny, nz):
    check(lib().vk_p2s(fp, fpt, I(nx), I(ny), I(nz)), "P2S")


def s2p(fpt, fp, nx, ny, nz):
    check(lib().vk_s2p(fpt, fp, I(nx), I(ny), I(nz)), "S2P")


def spn(nx, ny, nz, ac, kxz, phit):
    check(lib().vk_spn(I(nx), I(ny), I(nz), ac, kxz, phit), "SPN")


def sp0b(nx, ny, nz, ac, phit):
    check(lib().vk_sp0b(I(nx), I(ny), I(nz), ac, phit), "SP0B")


def phibc(nx, ny, nz, phi):
    check(lib().vk_phibc(I(nx), I(ny), I(nz), phi), "PHIBC")


def set_levelset(nx, ny, nz, dx, dx12, dy12, dz12, phi0, phi, x, y, z, cellvalue):
    check(lib().vk_set_levelset(I(nx), I(ny), I(nz), dx, dx12, dy12, dz12, phi0, phi, x, y, z,
                                cellvalue), "set_levelset")


def levelset_shapes(nx, ny, nz, x, y, z, xl, zl, shapes, phi0):
    """phi0 (device) of the union of ellipsoid records `shapes` ((n, 10):
    centre, semi-axes, quaternion), positive inside, halo 0
    (vk_levelset_shapes)."""
    import ctypes
    sh = np.ascontiguousarray(np.asarray(shapes, dtype=np.float64).reshape(-1, 10))
    check(lib().vk_levelset_shapes(I(nx), I(ny), I(nz), x, y, z, D(xl), D(zl), I(len(sh)),
                                   sh.ctypes.data_as(ctypes.c_void_p), phi0),
          "levelset_shapes")


def device_phi0(cfg: G.Config, g: G.Grid, state: dict | None = None) -> DeviceArray:
    """phi0 of a configuration on the device: generated there from its
    ellipsoids, or uploaded from the host state's phi0."""
    if cfg.ellipsoids is not None:
        phi0 = DeviceArray(g.shape)
        xyz = [DeviceArray.from_host(getattr(g, n)) for n in ("x", "y", "z")]
        levelset_shapes(g.nx, g.ny, g.nz, *xyz, g.xl, g.zl, cfg.ellipsoids, phi0)
        synchronize()
        for a in xyz:
            a.free()
        return phi0
    st = state if state is not None else G.initial_state(cfg, g)
    return DeviceArray.from_host(st["phi0"])


def set_adj_convect(u, u_v, v, v_v, w, w_v, re, dt, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12,
                    u_final, v_final, w_final):
    check(lib().vk_set_adj_convect(u, u_v, v, v_v, w, w_v, D(re), D(dt), I(nx), I(ny), I(nz),
                                   dx, dy, dz, dx12, dy12, dz12, u_final, v_final, w_final),
          "set_adj_convect")


def set_adj_convection_for_scalar(u, v, w, re, dt, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12,
                                  t, u_vt, rhocp, u_final, v_final, w_final):
    check(lib().vk_set_adj_convection_for_scalar(u, v, w, D(re), D(dt), I(nx), I(ny), I(nz),
                                                 dx, dy, dz, dx12, dy12, dz12, t, u_vt, rhocp,
                                                 u_final, v_final, w_final),
          "set_adj_convection_for_scalar")


def set_adj_ns_scalar(dt, nx, ny, nz, dx12, dy12, dz12, t, u_s, v_s, w_s, u_final, v_final,
                      w_final, t_final):
    check(lib().vk_set_adj_ns_scalar(D(dt), I(nx), I(ny), I(nz), dx12, dy12, dz12, t, u_s, v_s,
                                     w_s, u_final, v_final, w_final, t_final),
          "set_adj_NS_scalar")


def extent(a) -> int:
    """Bytes from ``a`` (a DeviceArray, a device pointer or a registered
    host array) to the end of its allocation."""
    import ctypes
    n = ctypes.c_size_t(0)
    check(lib().vk_extent(a, ctypes.byref(n)), "extent")
    return n.value


def field_absmax(f, nx, ny, nz) -> float:
    import ctypes
    r = ctypes.c_double(0.0)
    check(lib().vk_field_absmax(f, I(nx), I(ny), I(nz), ctypes.byref(r)), "field_absmax")
    return r.value


def check_courant(dt, dx, dy, dz, nx, ny, nz, u, v, w) -> float:
    """CHECK_COURANT (nse_dominik.f:2519): the largest signed u dt/dx,
    v dt/dy, w dt/dz over the interior, from 0 (the reference prints it)."""
    import ctypes
    r = ctypes.c_double(0.0)
    check(lib().vk_check_courant(D(dt), dx, dy, dz, I(nx), I(ny), I(nz), u, v, w,
                                 ctypes.byref(r)), "check_courant")
    return r.value


class MonitorState:
    """MONITOR's variables kept between calls by program main (u_int_before,
    v_int_before, w_int_before, T_b) and the ones it sets."""

    NAMES = ("dif", "t_all", "t_dif", "vel_dif", "u_dif", "v_dif", "w_dif", "u_int_before",
             "v_int_before", "w_int_before", "t_b")

    def __init__(self):
        import ctypes
        self._c = {n: ctypes.c_double(0.0) for n in self.NAMES}
        self.loc = (0, 0, 0)

    def __getitem__(self, name):
        return self._c[name].value


def monitor(nx, ny, nz, l, t, dx, dy, dz, t_before, u, v, w, dt, state: MonitorState,
            int_report=0) -> MonitorState:
    """MONITOR (nse_dominik.f:2797) on device fields; ``state`` carries the
    scalars between calls as the reference's host variables do."""
    import ctypes
    c = state._c
    loc = (ctypes.c_int * 3)()
    check(lib().vk_monitor(I(nx), I(ny), I(nz), I(l), t, dx, dy, dz, ctypes.byref(c["dif"]),
                           ctypes.byref(c["t_all"]), ctypes.byref(c["t_dif"]), t_before, u, v,
                           w, ctypes.byref(c["vel_dif"]), ctypes.byref(c["u_dif"]),
                           ctypes.byref(c["v_dif"]), ctypes.byref(c["w_dif"]),
                           ctypes.byref(c["u_int_before"]), ctypes.byref(c["v_int_before"]),
                           ctypes.byref(c["w_int_before"]), ctypes.byref(c["t_b"]), D(dt),
                           I(int_report), loc), "monitor")
    state.loc = tuple(loc)
    return state


def check_convergence(y, yl, u, l, dt, re, nx, ny, nz) -> float:
    """CHECK_CONVERGENCE (nse_dominik.f:2875): the error sum over the plane
    i = nx against the Poiseuille profile (the reference appends it to
    timely_convergence.dat with l*dt)."""
    import ctypes
    r = ctypes.c_double(0.0)
    check(lib().vk_check_convergence(y, D(yl), u, I(l), D(dt), D(re), I(nx), I(ny), I(nz),
                                     ctypes.byref(r)), "check_convergence")
    return r.value


def readgrid3d(nx, ny, nz, phi0: np.ndarray) -> np.ndarray:
    """READGRID3D (set_levelset.f:202-214): the interior of the halo array
    phi0 (float64, C-contiguous, shape (nz+2, ny+2, nx+2)) from
    structure.dat in the working directory, in place."""
    import ctypes
    if phi0.dtype != np.float64 or not phi0.flags.c_contiguous or \
            phi0.shape != (nz + 2, ny + 2, nx + 2):
        raise ValueError("phi0 must be a C-contiguous float64 (nz+2, ny+2, nx+2) array")
    check(lib().vk_readgrid3d(I(nx), I(ny), I(nz), phi0.ctypes.data_as(ctypes.c_void_p)),
          "readgrid3d")
    return phi0


def _host(a, shape, name):
    """Pointer to a host array the field-file routines read or write in
    full: the library moves shape's product of doubles, so a wrongly sized
    array would be over-read or overwritten."""
    import ctypes
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags.c_contiguous:
        raise ValueError(f"{name}: field files take C-contiguous float64 host arrays")
    if a.shape != tuple(shape):
        raise ValueError(f"{name}: shape {a.shape}, need {tuple(shape)}")
    return a.ctypes.data_as(ctypes.c_void_p)


def _coords(g: G.Grid):
    return (_host(g.x, (g.nx + 2,), "x"), _host(g.y, (g.ny + 2,), "y"),
            _host(g.z, (g.nz + 2,), "z"))


def outfield(g: G.Grid, tid: float, fields: dict) -> None:
    """OUTFIELD (nse_dominik.f:2726): field.u and ./nseoutput2/grid.u in the
    working directory from host arrays u, v, w, p, t, phi, u_ta, v_ta, w_ta
    (halo shape (nz+2, ny+2, nx+2))."""
    f = [_host(fields[k], g.shape, k) for k in ("u", "v", "w", "p")]
    r = [_host(fields[k], g.shape, k) for k in ("t", "phi", "u_ta", "v_ta", "w_ta")]
    check(lib().vk_outfield(*_coords(g), *f, I(g.nx), I(g.ny), I(g.nz), D(tid), *r),
          "outfield")


def output(g: G.Grid, fields: dict, cellvalue: np.ndarray) -> None:
    """OUTPUT (nse_dominik.f:2606): the nseoutput2/ files and
    temperature_check.dat in the working directory, from host arrays u, v,
    w, p, t, u_ta, v_ta, w_ta, phi, divu (halo shape) and cellvalue
    (4, nz, ny, nx)."""
    f = {k: _host(fields[k], g.shape, k) for k in ("u", "v", "w", "p", "t", "u_ta", "v_ta",
                                                   "w_ta", "phi", "divu")}
    cv = _host(cellvalue, (4, g.nz, g.ny, g.nx), "cellvalue")
    check(lib().vk_output(f["u"], f["v"], f["w"], f["p"], *_coords(g), I(g.nx), I(g.ny),
                          I(g.nz), f["t"], cv, f["u_ta"], f["v_ta"], f["w_ta"], f["phi"],
                          f["divu"]), "output")


def readfield(g: G.Grid, fields: dict) -> dict:
    """READFIELD (nse_dominik.f:2769): ./nseoutput2/field.u into the host
    arrays u, v, w, p, t, phi, u_ta, v_ta, w_ta (halo shape), in place."""
    f = [_host(fields[k], g.shape, k) for k in ("u", "v", "w", "p")]
    r = [_host(fields[k], g.shape, k) for k in ("t", "phi", "u_ta", "v_ta", "w_ta")]
    check(lib().vk_readfield(*f, I(g.nx), I(g.ny), I(g.nz), *r), "readfield")
    return fields


def _ptr(a):
    """A host numpy array's or a DeviceArray's address for the file routines."""
    import ctypes
    if isinstance(a, DeviceArray):
        return a.ptr
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags.c_contiguous:
        raise ValueError("file routines take C-contiguous float64 host arrays or DeviceArrays")
    return a.ctypes.data_as(ctypes.c_void_p)


def _checked(a, shape, name):
    if tuple(a.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(a.shape)}, need {tuple(shape)}")
    if isinstance(a, DeviceArray) and a.dtype != np.float64:
        raise ValueError(f"{name}: float64 arrays only")
    return _ptr(a)


def vtk_levelset(g: G.Grid, phi, path: str | None = None) -> None:
    """levelset.vtk (set_levelset.f:131-156) from phi (halo shape, host or
    device), byte for byte as the reference writes it."""
    p = path.encode() if path else None
    check(lib().vk_vtk_levelset(I(g.nx), I(g.ny), I(g.nz), *_coords(g),
                                _checked(phi, g.shape, "phi"), p), "vtk_levelset")


def vtk_cellvalue(g: G.Grid, cellvalue, path: str | None = None) -> None:
    """cellvalue.vtk (nse_dominik.f:208-235; cellvalue1.vtk with that path,
    set_levelset.f:171-198) from cellvalue (4, nz, ny, nx), host or device."""
    p = path.encode() if path else None
    check(lib().vk_vtk_cellvalue(I(g.nx), I(g.ny), I(g.nz), *_coords(g),
                                 _checked(cellvalue, (4, g.nz, g.ny, g.nx), "cellvalue"), p),
          "vtk_cellvalue")


RESTART_FIELDS = ("u", "v", "w", "p", "t", "u_pre", "v_pre", "w_pre", "t_pre", "u_ta", "v_ta",
                  "w_ta", "rhocp", "thc")


def write_restart(path: str, g: G.Grid, l: int, fields: dict, cellvalue) -> None:
    """The restart record (vk_write_restart): the fields of RESTART_FIELDS
    (halo shape, float64, host or device), cellvalue and the step l."""
    ptrs = [_checked(fields[k], g.shape, k) for k in RESTART_FIELDS]
    check(lib().vk_write_restart(path.encode(), I(g.nx), I(g.ny), I(g.nz), I(l), *ptrs,
                                 _checked(cellvalue, (4, g.nz, g.ny, g.nx), "cellvalue")),
          "write_restart")


def read_restart(path: str, g: G.Grid, fields: dict, cellvalue) -> int:
    """Reads a restart record into the arrays (vk_read_restart) and returns
    the step l it was written after."""
    import ctypes
    ptrs = [_checked(fields[k], g.shape, k) for k in RESTART_FIELDS]
    l = ctypes.c_int(0)
    check(lib().vk_read_restart(path.encode(), I(g.nx), I(g.ny), I(g.nz), ctypes.byref(l),
                                *ptrs, _checked(cellvalue, (4, g.nz, g.ny, g.nx), "cellvalue")),
          "read_restart")
    return l.value


def levelset_cellvalue(g: G.Grid, phi0: np.ndarray) -> np.ndarray:
    """cellvalue (4, nz, ny, nx) from set_levelset on the device (slot 1
    filled, slots 2-4 zero until a time step derives them)."""
    d = {n: DeviceArray.from_host(getattr(g, n)) for n in ("dx", "dx12", "dy12", "dz12",
                                                           "x", "y", "z")}
    p0 = phi0 if isinstance(phi0, DeviceArray) else DeviceArray.from_host(phi0)
    phi = DeviceArray(g.shape)
    cv = DeviceArray((4, g.nz, g.ny, g.nx))
    cv.fill_zero()
    set_levelset(g.nx, g.ny, g.nz, d["dx"], d["dx12"], d["dy12"], d["dz12"], p0, phi, d["x"],
                 d["y"], d["z"], cv)
    return cv.download()


# ------------------------------------------------------------- adjoint

ADJ_FIELDS = ("u", "v", "w", "p", "t", "u_pre", "v_pre", "w_pre", "t_pre", "u_ta", "v_ta",
              "w_ta")


class AdjointProblem:
    """An initial state, a level set and a run length for the wall-heat-flux
    objective and its adjoint gradient with respect to cellvalue(...,1)
    (vk_objective, vk_adjoint_gradient); every array resident in HBM."""

    def __init__(self, cfg: G.Config, g: G.Grid, state: dict, cellvalue: np.ndarray,
                 nsteps: int):
        init()
        self.cfg, self.g, self.nsteps = cfg, g, nsteps
        self.d = {}
        for n in ADJ_FIELDS + ("rhocp", "thc"):
            if state.get(n) is None:  # a lean state: the field starts at 0
                self.d[n] = DeviceArray(g.shape)
                self.d[n].fill_zero()
            else:
                self.d[n] = DeviceArray.from_host(state[n])
        for n in ("dx", "dy", "dz", "dx12", "dy12", "dz12"):
            self.d[n] = DeviceArray.from_host(getattr(g, n))
        self.cv = DeviceArray.from_host(np.asarray(cellvalue, dtype=np.float64))

    def _args(self, cellvalue):
        d, g, c = self.d, self.g, self.cfg
        return ([d[n] for n in ADJ_FIELDS] + [cellvalue, d["rhocp"], d["thc"]] +
                [d[n] for n in ("dx", "dy", "dz", "dx12", "dy12", "dz12")] +
                [I(g.nx), I(g.ny), I(g.nz), D(c.dt), D(c.re), D(c.sc), D(c.eta), D(g.xl),
                 D(g.zl), I(self.nsteps)])

    def objective(self, cellvalue=None) -> float:
        """J for the stored level set, or for `cellvalue` (host array, all
        four slots; slots 2-4 are recomputed at step 1)."""
        import ctypes
        cv = self.cv
        if cellvalue is not None:
            cv = DeviceArray.from_host(np.asarray(cellvalue, dtype=np.float64))
        out = ctypes.c_double(0.0)
        check(lib().vk_objective(*self._args(cv), ctypes.byref(out)), "objective")
        return out.value

    def gradient_revolve(self, snapshots: int = -1, mixed: bool = False):
        """(J, dJ/dcellvalue(...,1), stats) by binomial checkpointing with
        `snapshots` state slots (-1: as many as fit), about the FP64 or the
        mixed-precision trajectory (vk_adjoint_gradient_revolve).  stats:
        snapshots, forward steps run, recomputed steps, device bytes held."""
        import ctypes
        g = self.g
        grad = DeviceArray((g.nz, g.ny, g.nx))
        out = ctypes.c_double(0.0)
        stats = (ctypes.c_double * 4)()
        check(lib().vk_adjoint_gradient_revolve(*self._args(self.cv), I(snapshots),
                                                I(1 if mixed else 0), ctypes.byref(out), grad,
                                                stats), "adjoint_gradient_revolve")
        g_host = grad.download()
        grad.free()
        st = {"snapshots": int(stats[0]), "forward_steps": int(stats[1]),
              "recomputed_steps": int(stats[2]), "device_bytes": int(stats[3])}
        return out.value, g_host, st

    def gradient(self, checkpoint_every: int = 20):
        """(J, dJ/dcellvalue(...,1)) with the gradient as an (nz, ny, nx) array."""
        import ctypes
        g = self.g
        grad = DeviceArray((g.nz, g.ny, g.nx))
        out = ctypes.c_double(0.0)
        check(lib().vk_adjoint_gradient(*self._args(self.cv), I(checkpoint_every),
                                        ctypes.byref(out), grad), "adjoint_gradient")
        return out.value, grad.download()


# ------------------------------------------------------------------ runs

STEP_FIELDS = ("u", "v", "w", "p", "t", "u_pre", "v_pre", "w_pre", "t_pre",
               "u_ta", "v_ta", "w_ta", "u_temp", "v_temp", "w_temp", "divu",
               "rhocp", "thc", "u_ep2", "v_ep2", "w_ep2")


# fields the mixed-precision step keeps in float (with cellvalue)
MIXED_F32 = ("t", "u_pre", "v_pre", "w_pre", "t_pre", "u_ta", "v_ta", "w_ta", "rhocp", "thc")


class ChannelRun:
    """A configuration resident in HBM, advanced by vk_time_step
    (``precision="fp64"``) or vk_time_step_mixed (``precision="mixed"``).

    ``state`` is the dict of initial_state(); the level set is either
    given (``cellvalue``) or computed on the device by set_levelset.  In
    the mixed form the fields of MIXED_F32 and cellvalue are held in
    float, the rest in double; the MKP splitting-error diagnostic is
    not produced.
    """

    def __init__(self, cfg: G.Config, g: G.Grid, state: dict, cellvalue=None,
                 diagnostics=True, precision: str = "fp64"):
        init()
        if precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', not {precision!r}")
        self.precision = precision
        mixed = precision == "mixed"
        self.cfg, self.g = cfg, g
        self.d = {}
        for n in STEP_FIELDS:
            if mixed and n in ("u_ep2", "v_ep2", "w_ep2"):
                continue
            dtype = np.float32 if mixed and n in MIXED_F32 else np.float64
            a = state.get(n)
            if a is None:  # a lean state: the field starts at 0
                self.d[n] = DeviceArray(g.shape, dtype)
                self.d[n].fill_zero()
            else:
                self.d[n] = DeviceArray.from_host(np.asarray(a, dtype=dtype))
        for n in ("dx", "dy", "dz", "dx12", "dy12", "dz12", "x", "y", "z"):
            self.d[n] = DeviceArray.from_host(getattr(g, n))
        if cellvalue is None:
            self.d["phi0"] = device_phi0(cfg, g, state)
            self.d["phi"] = DeviceArray(g.shape)
            self.d["cellvalue"] = DeviceArray((4, g.nz, g.ny, g.nx))
            d = self.d
            set_levelset(g.nx, g.ny, g.nz, d["dx"], d["dx12"], d["dy12"], d["dz12"], d["phi0"],
                         d["phi"], d["x"], d["y"], d["z"], d["cellvalue"])
            if mixed:
                cv = self.d["cellvalue"].download()
                self.d["cellvalue"].free()
                self.d["cellvalue"] = DeviceArray.from_host(cv.astype(np.float32))
        else:
            cv = np.asarray(cellvalue, dtype=np.float32 if mixed else np.float64)
            self.d["cellvalue"] = DeviceArray.from_host(cv)
        self.diagnostics = diagnostics and not mixed
        self.l = 0
        self._call = None

    def _step_call(self):
        """The step's entry point and its arguments, converted once: the
        arrays do not move while the run lives, and the step counter is a
        c_int updated in place, so a step costs one call and no conversions
        (the mixed step waits for its projection each step, and whatever the
        host does between steps the GPU waits for)."""
        if self._call is None:
            import ctypes
            d, g, c = self.d, self.g, self.cfg
            fields = ["u", "v", "w", "p", "t", "u_pre", "v_pre", "w_pre", "t_pre", "u_ta",
                      "v_ta", "w_ta", "u_temp", "v_temp", "w_temp", "divu", "cellvalue",
                      "rhocp", "thc"]
            ptrs = [d[k]._as_parameter_ for k in fields]
            if self.precision == "fp64":
                ptrs += [d[k]._as_parameter_ if self.diagnostics else None
                         for k in ("u_ep2", "v_ep2", "w_ep2")]
            ptrs += [d[k]._as_parameter_ for k in ("dx", "dy", "dz", "dx12", "dy12", "dz12")]
            self._lc = ctypes.c_int(0)
            args = tuple(ptrs) + (I(g.nx), I(g.ny), I(g.nz), D(c.dt), D(c.re), D(c.sc),
                                  D(c.eta), D(g.xl), D(g.zl), ctypes.byref(self._lc))
            fn = lib().vk_time_step_mixed if self.precision == "mixed" else lib().vk_time_step
            self._call = (fn, args)
        return self._call

    def step(self, n: int = 1) -> None:
        fn, args = self._step_call()
        what = "mixed time step" if self.precision == "mixed" else "time step"
        for _ in range(n):
            self.l += 1
            self._lc.value = self.l
            rc = fn(*args)
            if rc:
                check(rc, f"{what} {self.l}")

    def get(self, name: str) -> np.ndarray:
        """A field downloaded to the host, as float64 whatever its storage."""
        return self.d[name].download().astype(np.float64, copy=False)

    def save(self, path: str) -> None:
        """Writes the run's restart record (everything the next step reads,
        and the step counter).  The mixed form's float fields are written
        as doubles (exactly), so the record is the same format."""
        names = RESTART_FIELDS + ("cellvalue",)
        if self.precision == "fp64":
            arrs = {k: self.d[k] for k in names}
        else:
            arrs = {k: self.d[k].download().astype(np.float64) for k in names}
        write_restart(path, self.g, self.l, arrs, arrs["cellvalue"])

    def restore(self, path: str) -> None:
        """Continues from a restart record written by save() (or by any
        caller of vk_write_restart) for the same grid."""
        names = RESTART_FIELDS + ("cellvalue",)
        if self.precision == "fp64":
            self.l = read_restart(path, self.g, {k: self.d[k] for k in names}, self.d["cellvalue"])
            return
        host = {k: np.zeros(self.d[k].shape) for k in names}
        self.l = read_restart(path, self.g, host, host["cellvalue"])
        for k in names:
            self.d[k].upload(host[k].astype(self.d[k].dtype))

    def free(self) -> None:
        """Releases every device array of the run now (rather than when
        the Python objects are collected)."""
        for a in self.d.values():
            a.free()
        self.d = {}
        self._call = None
