"""PHIBC (mkp.f:293-320) on its own: the halo of a pressure-like field.

The reference's PHIBC writes most of the halo from interior values, reads
some halo cells before it has set them, and leaves a few corner halos
unset.  Starting from NaN halos, the cells the oracle (pinned bit for bit
against the reference build) leaves finite are exactly those the
reference sets from the interior; there the library must agree bit for
bit.  Everywhere else the library defines the halo: x and z periodic, y
zero-gradient, corners by composition, which is checked against numpy.
"""
import numpy as np
import pytest

from oracle_bind import P, oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [(8, 9, 8), (12, 7, 10), (48, 33, 40)])
def test_phibc_halos(n):
    from vtkmakecompare_amd import api as A
    A.init()
    nx, ny, nz = n
    rng = np.random.default_rng(nx * ny + nz)
    phi = np.full((nz + 2, ny + 2, nx + 2), np.nan)
    phi[1:-1, 1:-1, 1:-1] = rng.normal(size=(nz, ny, nx))
    want = phi.copy()
    oracle().o_phibc(nx, ny, nz, P(want))
    d = A.DeviceArray.from_host(phi)
    A.phibc(nx, ny, nz, d)
    got = d.download()
    written = np.isfinite(want)
    assert written.sum() > phi[1:-1, 1:-1, 1:-1].size  # the reference writes halos
    np.testing.assert_array_equal(got[written], want[written])
    # the library's full halo: periodic x and z, zero-gradient y
    full = phi.copy()
    full[1:-1, 1:-1, 0] = full[1:-1, 1:-1, nx]
    full[1:-1, 1:-1, nx + 1] = full[1:-1, 1:-1, 1]
    full[0, 1:-1, :] = full[nz, 1:-1, :]
    full[nz + 1, 1:-1, :] = full[1, 1:-1, :]
    full[:, 0, :] = full[:, 1, :]
    full[:, ny + 1, :] = full[:, ny, :]
    assert np.all(np.isfinite(got))
    np.testing.assert_array_equal(got, full)
