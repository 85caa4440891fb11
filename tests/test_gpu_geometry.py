"""The analytic geometry front-end on the device (vk_levelset_shapes): phi0
of a union of rotated ellipsoids against the oracle's restatement (whose
closed-form pins are in test_oracle_math.py), including centres near the
periodic x and z boundaries, and the argument refusals.
"""
import numpy as np
import pytest

from oracle_bind import oracle_shapes
from vtkmakecompare_amd import grid as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from vtkmakecompare_amd import api as A
    A.init()
    return A


def _device_phi0(A, g, recs):
    xyz = [A.DeviceArray.from_host(getattr(g, n)) for n in ("x", "y", "z")]
    phi0 = A.DeviceArray(g.shape)
    A.levelset_shapes(g.nx, g.ny, g.nz, *xyz, g.xl, g.zl, recs, phi0)
    return phi0.download()


@pytest.mark.parametrize("n,count", [((48, 33, 40), 7), ((96, 65, 96), 64), ((20, 9, 12), 0)])
def test_shapes_match_oracle(api, n, count):
    g = G.channel_grid(*n)
    recs = G.random_ellipsoids(0x5EED0301 + count, count, g.xl, g.zl, 0.1, 0.5)
    if count:
        recs[0, 0] = 0.02  # wraps in x
        recs[1, 2] = g.zl - 0.03  # wraps in z
    got = _device_phi0(api, g, recs)
    want = oracle_shapes(g, recs)
    err = float(np.max(np.abs(got - want)))
    print(f"levelset_shapes {n}, {count} shapes: max |gpu - oracle| {err:.2e}")
    assert err <= 4e-16 * max(float(np.max(np.abs(want))), 1.0)
    assert np.all(got[0] == 0.0) and np.all(got[:, 0] == 0.0) and np.all(got[:, :, 0] == 0.0)


def test_shapes_feed_set_levelset(api):
    """The generated phi0 goes through set_levelset like a structure.dat
    one: cellvalue(...,1) in [0, 1], 1 deep inside a shape, 0 far away."""
    g = G.channel_grid(48, 33, 40)
    # semi-axes of several cells (dx = 0.13 here), so the reinitialised
    # distance at the centre stays above dx, where cellvalue is exactly 1
    recs = np.array([[np.pi, 0.0, 0.5 * np.pi, 1.2, 0.7, 0.6, 1.0, 0.2, -0.3, 0.1]])
    phi0 = _device_phi0(api, g, recs)
    cv = api.levelset_cellvalue(g, phi0)
    assert cv[0].min() >= 0.0 and cv[0].max() <= 1.0
    i = int(np.argmin(np.abs(g.x[1:-1] - np.pi)))
    k = int(np.argmin(np.abs(g.z[1:-1] - 0.5 * np.pi)))
    j = int(np.argmin(np.abs(g.y[1:-1])))
    assert cv[0, k, j, i] == 1.0 and cv[0, 0, 0, 0] == 0.0


def test_shapes_refusals(api):
    g = G.channel_grid(8, 9, 8)
    xyz = [api.DeviceArray.from_host(getattr(g, n)) for n in ("x", "y", "z")]
    phi0 = api.DeviceArray(g.shape)
    bad_axis = np.array([[1.0, 0.0, 1.0, 0.2, 0.0, 0.2, 1.0, 0.0, 0.0, 0.0]])
    zero_q = np.array([[1.0, 0.0, 1.0, 0.2, 0.2, 0.2, 0.0, 0.0, 0.0, 0.0]])
    for recs in (bad_axis, zero_q):
        with pytest.raises(api.VKError) as e:
            api.levelset_shapes(g.nx, g.ny, g.nz, *xyz, g.xl, g.zl, recs, phi0)
        assert e.value.status == 5  # VK_ERR_ARGUMENT
    short = api.DeviceArray((g.nz + 2, g.ny + 2, g.nx + 1))
    with pytest.raises(api.VKError):
        api.levelset_shapes(g.nx, g.ny, g.nz, *xyz, g.xl, g.zl, bad_axis[:0], short)
