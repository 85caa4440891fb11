"""Loader of the gfx950 library libvtkgpu.so (the C ABI of include/vtkgpu.h).

The library is built in-tree by __graft_entry__.build() (or
``make -C vtkmakecompare_amd/csrc``).  There is no fallback: if it is
missing or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvtkgpu.so")

# every exported symbol of include/vtkgpu.h
EXPORTS = [
    "vk_init", "vk_finalize", "vk_last_error", "vk_register", "vk_unregister",
    "vk_upload", "vk_download", "vk_synchronize", "vk_malloc", "vk_free",
    "vk_memcpy_htod", "vk_memcpy_dtoh", "vk_memset", "vk_extent", "vk_guard_check",
    "vk_set_bc", "vk_set_convect", "vk_set_convection_for_scalar",
    "vk_set_conduction", "vk_set_diffusion_for_scalar",
    "vk_set_immersed_boundary", "vk_set_explicit",
    "vk_set_explicit_for_tempreture", "vk_t_divide", "vk_solve_u_temp",
    "vk_solve_t", "vk_ctridiag_batched", "vk_tdma_batched", "vk_cyclic_batched",
    "vk_implicit_solver_batched", "vk_implicit_uy_batched", "vk_implicit_vy_batched",
    "vk_implicit_tx_batched", "vk_implicit_ty_batched", "vk_linebyline", "vk_solve_p",
    "vk_mkp", "vk_p2s", "vk_s2p", "vk_spn", "vk_sp0b",
    "vk_phibc", "vk_set_levelset", "vk_levelset_shapes", "vk_set_adj_convect",
    "vk_set_adj_convection_for_scalar", "vk_set_adj_ns_scalar",
    "vk_time_step", "vk_time_step_mixed", "vk_objective", "vk_adjoint_gradient",
    "vk_adjoint_gradient_revolve",
    "vk_field_absmax", "vk_field_copy", "vk_check_courant", "vk_check_courant_noalloc",
    "vk_monitor",
    "vk_check_convergence", "vk_readgrid3d", "vk_outfield", "vk_readfield",
    "vk_output", "vk_vtk_levelset", "vk_vtk_cellvalue", "vk_write_restart",
    "vk_read_restart",
    "vk_profile_begin", "vk_profile_end", "vk_profile_count", "vk_profile_entry",
]


class VKError(RuntimeError):
    """A non-zero status of the library; ``status`` is the VK_ERR_* code."""

    def __init__(self, msg: str, status: int | None = None):
        super().__init__(msg)
        self.status = status


_lib = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises VKError if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VKError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name in EXPORTS:
            f = getattr(L, name)
            f.restype = ctypes.c_int
        L.vk_last_error.restype = ctypes.c_char_p
        L.vk_last_error.argtypes = []
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().vk_last_error().decode(errors="replace")
        raise VKError(f"{what}: status {rc}: {msg}", rc)


def I(v: int):
    return ctypes.byref(ctypes.c_int(int(v)))


def D(v: float):
    return ctypes.byref(ctypes.c_double(float(v)))


def S(v: int):
    return ctypes.byref(ctypes.c_size_t(int(v)))
