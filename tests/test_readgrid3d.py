"""READGRID3D (set_levelset.f:202-214) against the reference's own compiled
routine, bit for bit, on structure.dat files written the ways a
list-directed read accepts: plain and E/D exponents, comma separators,
extra values after the fourth (skipped with the rest of the record), and
an item continued on the next record.  The library reads into host memory
here (no GPU is touched for an unregistered array).
"""
import os

import numpy as np
import pytest

from oracle_bind import I, P, ref
from vtkmakecompare_amd import api as A
from vtkmakecompare_amd import grid as G


def fortran_d(v):
    """v with a D exponent, as Fortran writes double precision."""
    return f"{v:.16E}".replace("E", "D")


def write_structure(path, g, phi, continued=True):
    """structure.dat in every form the reader takes; with `continued`,
    some items span two records (the serial reader's case), otherwise
    every item is one record (the parallel reader's case)."""
    rng = np.random.default_rng(17)
    lines = []
    for k in range(1, g.nz + 1):
        for j in range(1, g.ny + 1):
            for i in range(1, g.nx + 1):
                v = float(phi[k, j, i])
                form = rng.integers(0, 5 if continued else 4)
                a, b, c = float(g.x[i]), float(g.y[j]), float(g.z[k])
                if form == 0:
                    lines.append(f"{a!r} {b!r} {c!r} {v!r}")
                elif form == 1:
                    lines.append(f"  {fortran_d(a)}  {fortran_d(b)}  {fortran_d(c)}  {fortran_d(v)}")
                elif form == 2:
                    lines.append(f"{a:.17g}, {b:.17g},{c:.17g} ,{v:.17e}")
                elif form == 3:
                    lines.append(f"{a!r} {b!r} {c!r} {v!r} 99.0 extra")
                else:
                    lines.append(f"{a!r} {b!r}")
                    lines.append(f"{c!r} {v!r}")
    path.write_text("\n".join(lines) + "\n")


@pytest.mark.skipif(ref() is None, reason="reference library not built")
@pytest.mark.parametrize("n,continued", [((6, 5, 4), True), ((9, 7, 3), True),
                                         ((40, 30, 20), False), ((40, 30, 20), True)])
def test_readgrid3d_matches_reference(tmp_path, n, continued):
    g = G.channel_grid(*n)
    phi = np.random.default_rng(3).normal(size=g.shape) * 10.0 ** np.random.default_rng(4).integers(
        -8, 8, size=g.shape)
    write_structure(tmp_path / "structure.dat", g, phi, continued)
    got = np.full(g.shape, 7.0)
    want = np.full(g.shape, 7.0)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        A.readgrid3d(g.nx, g.ny, g.nz, got)
        ref().readgrid3d_(I(g.nx), I(g.ny), I(g.nz), P(want))
    finally:
        os.chdir(cwd)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(got[1:-1, 1:-1, 1:-1], phi[1:-1, 1:-1, 1:-1])
    assert np.all(got[0] == 7.0) and np.all(got[-1] == 7.0)  # the halo is left alone


def test_readgrid3d_refusals(tmp_path):
    g = G.channel_grid(4, 3, 2)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        with pytest.raises(A.VKError):  # no structure.dat
            A.readgrid3d(g.nx, g.ny, g.nz, np.zeros(g.shape))
        (tmp_path / "structure.dat").write_text("1 2 3 4\n" * 5)
        with pytest.raises(A.VKError):  # too few records
            A.readgrid3d(g.nx, g.ny, g.nz, np.zeros(g.shape))
        (tmp_path / "structure.dat").write_text("1 2 3 2*4\n" * 24)
        with pytest.raises(A.VKError):  # repeat count
            A.readgrid3d(g.nx, g.ny, g.nz, np.zeros(g.shape))
        # a leading comma is a null value for a (list-directed input would
        # then give phi0 = 3 from ',1,2,3,4'), refused like any other null
        (tmp_path / "structure.dat").write_text("1 2 3 4\n" * 23 + ",1,2,3,4\n")
        with pytest.raises(A.VKError):
            A.readgrid3d(g.nx, g.ny, g.nz, np.zeros(g.shape))
        (tmp_path / "structure.dat").write_text("1 2 3 4\n" * 23 + "  , 1 2 3 4\n")
        with pytest.raises(A.VKError):
            A.readgrid3d(g.nx, g.ny, g.nz, np.zeros(g.shape))
    finally:
        os.chdir(cwd)
