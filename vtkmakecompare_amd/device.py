"""Device-resident arrays on top of the library's own allocator.

A :class:`DeviceArray` is device memory (vk_malloc) with the shape and
layout of a reference array; upload/download are flat copies because the
device layout is the reference's host layout.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import S, check, lib

_initialised = False


def init(device: int = 0) -> None:
    global _initialised
    if not _initialised:
        check(lib().vk_init(ctypes.byref(ctypes.c_int(device))), "vk_init")
        _initialised = True


def synchronize() -> None:
    check(lib().vk_synchronize(), "vk_synchronize")


class DeviceArray:
    def __init__(self, shape, dtype=np.float64):
        init()
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().vk_malloc(ctypes.byref(p), S(self.nbytes)), "vk_malloc")
        self.ptr = p

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        check(lib().vk_memcpy_htod(self.ptr, a.ctypes.data_as(ctypes.c_void_p), S(self.nbytes)),
              "upload")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(lib().vk_memcpy_dtoh(out.ctypes.data_as(ctypes.c_void_p), self.ptr, S(self.nbytes)),
              "download")
        return out

    def fill_zero(self) -> None:
        check(lib().vk_memset(self.ptr, ctypes.byref(ctypes.c_int(0)), S(self.nbytes)), "memset")

    @property
    def _as_parameter_(self):
        return self.ptr

    def free(self) -> None:
        if self.ptr is not None and self.ptr.value:
            lib().vk_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
