"""ctypes bindings for the test-only checkers: the C oracle (oracle/) and
the reference routines compiled from the reference sources (oracle/_ref).

Only tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  The oracle's shared library is built by `make -C oracle` (or
__graft_entry__.build()); the reference build exists only where the
reference sources were present at build time.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libref.so")

_oracle = None
_ref = None


def P(a: np.ndarray):
    """Pointer to a contiguous float64 numpy array."""
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(ctypes.c_void_p)


def I(v: int):
    return ctypes.byref(ctypes.c_int(int(v)))


def D(v: float):
    return ctypes.byref(ctypes.c_double(float(v)))


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", ORACLE_DIR, "liboracle.so"], check=True,
                           stdout=subprocess.DEVNULL)
        lib = ctypes.CDLL(ORACLE_SO)
        _declare_oracle(lib)
        _oracle = lib
    return _oracle


def ref():
    """The reference routines (lazy binding: p2s_/s2p_ are unresolved, so
    mkp_ must never be called).  None when not built."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        _ref = ctypes.CDLL(REF_SO, mode=os.RTLD_LAZY)
    return _ref


def _declare_oracle(lib):
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    sig = {
        "o_set_c": [ci, ci, ci] + [vp] * 12 + [cd, cd, cd],
        "o_set_bc": [vp, vp, vp, vp, ci, ci, ci, vp],
        "o_set_convect": [vp] * 6 + [cd, ci, ci, ci] + [vp] * 6,
        "o_set_convection_for_scalar": [vp, vp, vp, cd, ci, ci, ci] + [vp] * 9,
        "o_set_conduction": [vp] * 6 + [cd, cd, ci, ci, ci] + [vp] * 6,
        "o_set_diffusion_for_scalar": [cd, cd, ci, ci, ci] + [vp] * 8 + [cd, vp, vp],
        "o_set_immersed_boundary": [cd, ci, ci, ci, cd] + [vp] * 11 + [ci, vp],
        "o_set_explicit": [vp] * 15 + [ci, ci, ci, ci],
        "o_set_explicit_for_tempreture": [vp] * 5 + [ci, ci, ci, ci],
        "o_t_divide": [vp, vp, vp, cd, cd, ci, ci, ci],
        "o_solve_u_temp": [vp, vp, vp, ci, ci, ci] + [vp] * 6 + [cd, cd, cd] + [vp] * 4,
        "o_solve_t": [vp, ci, ci, ci] + [vp] * 6 + [cd, cd, cd, vp, vp, vp, cd],
        "o_implicit_uy": [vp, ci, vp, vp, cd, cd, cd, cd, cd, vp],
        "o_implicit_vy": [vp, ci, vp, vp, cd, cd, cd, cd, cd, vp],
        "o_implicit_tx": [vp, ci, vp, vp, cd, cd, cd, vp, vp, cd, cd, cd, vp],
        "o_implicit_ty": [vp, ci, vp, vp, cd, cd, cd, vp, vp, cd, cd, vp, cd],
        "o_implicit_solver": [vp, ci, vp, vp, cd, cd, cd, cd, cd, vp],
        "o_tdma": [vp] * 5 + [ci],
        "o_cyclic": [vp, vp, vp, cd, cd, vp, vp, ci],
        "o_ctridiag": [vp, vp, ci],
        "o_mkp": [vp] * 4 + [cd] + [vp] * 6 + [cd, ci, ci, ci] + [vp] * 10 + [cd, cd],
        "o_p2s": [vp, vp, ci, ci, ci],
        "o_s2p": [vp, vp, ci, ci, ci],
        "o_spn": [ci, ci, ci, vp, vp, vp],
        "o_sp0b": [ci, ci, ci, vp, vp],
        "o_phibc": [ci, ci, ci, vp],
        "o_mkp_coefficients": [ci, ci, ci] + [vp] * 4 + [cd, cd, vp, vp],
        "o_fft": [vp, ci, ci, ci],
        "o_set_levelset": [ci, ci, ci] + [vp] * 10 + [ci],
        "o_levelset_shapes": [ci, ci, ci, vp, vp, vp, cd, cd, ci, vp, vp],
        "o_linebyline": [ci, ci, ci] + [vp] * 9 + [cd, vp],
        "o_solve_p": [vp] * 4 + [cd] + [vp] * 6 + [ci, ci, ci] + [vp] * 4 + [ci, vp],
        "o_set_adj_convect": [vp] * 6 + [cd, cd, ci, ci, ci] + [vp] * 9,
        "o_set_adj_convection_for_scalar": [vp, vp, vp, cd, cd, ci, ci, ci] + [vp] * 12,
        "o_set_adj_ns_scalar": [cd, ci, ci, ci] + [vp] * 11,
        "o_time_step": [vp, ci],
        "o_check_courant": [cd, vp, vp, vp, ci, ci, ci, vp, vp, vp, vp],
        "o_monitor": [ci, ci, ci, ci, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                      vp, vp, vp, vp, cd, vp],
        "o_check_convergence": [vp, cd, vp, cd, ci, ci, ci, vp],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = None


# ------------------------------------------------------------ the state

STATE_FIELDS = ["dx", "dy", "dz", "dx12", "dy12", "dz12",
                "u", "v", "w", "p", "t",
                "u_v", "v_v", "w_v", "u_d", "v_d", "w_d", "u_a", "v_a", "w_a",
                "u_temp", "v_temp", "w_temp", "u_pre", "v_pre", "w_pre",
                "u_ta", "v_ta", "w_ta", "u_l", "v_l", "w_l",
                "u_vt", "u_dt", "u_t", "t_pre", "t_l",
                "rhocp", "thc", "cellvalue",
                "u_ep2", "v_ep2", "w_ep2", "divu"]


class OState(ctypes.Structure):
    _fields_ = ([("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int)] +
                [(n, ctypes.c_double) for n in ("dt", "re", "sc", "eta", "xl", "zl")] +
                [(n, ctypes.c_void_p) for n in STATE_FIELDS])


class OracleRun:
    """The oracle's time loop over a numpy state dict (arrays owned here)."""

    def __init__(self, cfg, grid, state: dict):
        self.cfg, self.g = cfg, grid
        self.arrays = {}
        for n in STATE_FIELDS:
            if n in ("dx", "dy", "dz", "dx12", "dy12", "dz12"):
                self.arrays[n] = np.ascontiguousarray(getattr(grid, n), dtype=np.float64)
            else:
                self.arrays[n] = np.ascontiguousarray(state[n], dtype=np.float64).copy()
        s = OState()
        s.nx, s.ny, s.nz = grid.nx, grid.ny, grid.nz
        s.dt, s.re, s.sc, s.eta = cfg.dt, cfg.re, cfg.sc, cfg.eta
        s.xl, s.zl = grid.xl, grid.zl
        for n in STATE_FIELDS:
            setattr(s, n, self.arrays[n].ctypes.data)
        self.s = s
        self.l = 0

    def step(self, nsteps=1):
        lib = oracle()
        for _ in range(nsteps):
            self.l += 1
            lib.o_time_step(ctypes.byref(self.s), self.l)

    def __getitem__(self, name):
        return self.arrays[name]


def oracle_levelset(grid, phi0, iterations=2001):
    """cellvalue(...,1) and phi from the oracle's set_levelset."""
    lib = oracle()
    phi = np.zeros(grid.shape)
    cv = np.zeros((4, grid.nz, grid.ny, grid.nx))
    lib.o_set_levelset(grid.nx, grid.ny, grid.nz, P(grid.dx), P(grid.dx12),
                       P(grid.dy12), P(grid.dz12), P(np.ascontiguousarray(phi0)),
                       P(phi), P(grid.x), P(grid.y), P(grid.z), P(cv), iterations)
    return phi, cv


def oracle_shapes(grid, shapes):
    """phi0 of the analytic shape union from the oracle."""
    sh = np.ascontiguousarray(np.asarray(shapes, dtype=np.float64).reshape(-1, 10))
    phi0 = np.zeros(grid.shape)
    oracle().o_levelset_shapes(grid.nx, grid.ny, grid.nz, P(grid.x), P(grid.y), P(grid.z),
                               grid.xl, grid.zl, len(sh), P(sh), P(phi0))
    return phi0
