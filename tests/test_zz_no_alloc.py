"""vk_check_courant_noalloc, the entry point that allocates nothing,
against vk_check_courant through the C ABI: the same Courant number bit
for bit, on two calls in a row, the second on a smaller grid (so the
static scratch holds the first call's partials when the second starts),
and first of all on a fresh library where no workspace exists yet.
"""
import ctypes

import numpy as np
import pytest

from vtkmakecompare_amd import grid as G

pytestmark = pytest.mark.gpu


def _courant(lib, name, dt, m, g, f):
    out = ctypes.c_double(-1.0)
    I = lambda v: ctypes.byref(ctypes.c_int(v))  # noqa: E731
    rc = getattr(lib, name)(ctypes.byref(ctypes.c_double(dt)), m["dx"], m["dy"], m["dz"],
                            I(g.nx), I(g.ny), I(g.nz), f[0], f[1], f[2], ctypes.byref(out))
    assert rc == 0, (name, rc, lib.vk_last_error())
    return out.value


def _case(A, n, seed):
    g = G.channel_grid(*n)
    rng = np.random.default_rng(seed)
    m = {k: A.DeviceArray.from_host(getattr(g, k)) for k in ("dx", "dy", "dz")}
    f = [A.DeviceArray.from_host(rng.uniform(-2.0, 2.0, size=g.shape)) for _ in range(3)]
    return g, m, f


def test_noalloc_courant_matches_allocating_entry_point():
    from vtkmakecompare_amd import _lib
    from vtkmakecompare_amd import api as A
    A.init()
    lib = _lib.lib()
    dt = 1e-3
    # a grid of 1024 blocks of partials, then one of 6 blocks
    for n, seed in (((96, 49, 64), 1), ((16, 9, 10), 2)):
        g, m, f = _case(A, n, seed)
        a = _courant(lib, "vk_check_courant_noalloc", dt, m, g, f)
        b = _courant(lib, "vk_check_courant", dt, m, g, f)
        print(f"courant {n}: noalloc {a!r} allocating {b!r}")
        assert a == b and a > 0.0
        # again, now that the allocating entry point has run
        assert _courant(lib, "vk_check_courant_noalloc", dt, m, g, f) == b
