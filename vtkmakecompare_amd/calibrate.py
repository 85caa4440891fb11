"""A plain copy of known size, for scaling rocprofv3's memory counters.

    rocprofv3 --pmc FETCH_SIZE --kernel-trace -d DIR -o copy -- \
        python3 -m vtkmakecompare_amd.calibrate

FETCH_SIZE and WRITE_SIZE count the L2's requests to memory in units that
depend on the request size. A copy whose bytes are known (the library's
interior copy, ``vk_field_copy`` with ``interior_only = 1``, of a field far
larger than the L2 and the MALL) gives the factor that turns the counters
into bytes for the other kernels. Each launch reads and writes
``nx*ny*nz*8`` bytes; the copy runs ``--reps`` times, alternating direction.
"""
from __future__ import annotations

import argparse
import ctypes
import sys

from . import device
from ._lib import I, check, lib


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=510)
    ap.add_argument("--nz", type=int, default=1022)
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    device.init(0)
    cells = (a.nx + 2) * (a.ny + 2) * (a.nz + 2)
    x = device.DeviceArray((cells,))
    y = device.DeviceArray((cells,))
    x.fill_zero()
    y.fill_zero()
    device.synchronize()
    L = lib()
    for r in range(a.reps):
        src, dst = (x, y) if r % 2 == 0 else (y, x)
        check(L.vk_field_copy(src.ptr, dst.ptr, I(a.nx), I(a.ny), I(a.nz), I(1)), "field_copy")
    device.synchronize()
    print(f"calibrate: {a.reps} interior copies of {a.nx}x{a.ny}x{a.nz}, "
          f"{a.nx * a.ny * a.nz * 8} bytes read and written per launch", flush=True)
    x.free()
    y.free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
