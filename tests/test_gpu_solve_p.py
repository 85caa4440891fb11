"""The reference's iterative Poisson solver on the device (vk_linebyline,
vk_solve_p) against the oracle, whose linebyline and solve_p are pinned
bit for bit against the reference's compiled routines:

- linebyline: p after three iterations bit for bit (the wavefront order
  reproduces the serial sweep's Gauss-Seidel dependencies exactly), on
  the stretched channel grid and on a uniform one, with lines of unequal
  lengths in the three directions;
- solve_p to its 1e-5 residual on a uniform grid with a compatible
  right-hand side: the same iteration count, p bit for bit, the largest
  and the last residual bit for bit, the two integrals to rounding of
  their sums (a tree order here, the serial order there).
"""
import ctypes

import numpy as np
import pytest

from oracle_bind import P, oracle
from vtkmakecompare_amd import grid as G

pytestmark = pytest.mark.gpu


def _case(g, seed, scale=1.0, compatible=False):
    rng = np.random.default_rng(seed)
    ut, vt, wt = (scale * rng.uniform(-1, 1, size=g.shape) for _ in range(3))
    if compatible:
        ut[:, :, 0] = ut[:, :, g.nx] = 0.0
        vt[:, 0, :] = vt[:, g.ny, :] = 0.0
        wt[0] = wt[g.nz] = 0.0
    p = rng.uniform(-0.1, 0.1, size=g.shape)
    return ut, vt, wt, p


def _dev(A, g, *arrs):
    m = [A.DeviceArray.from_host(getattr(g, k)) for k in ("dx", "dy", "dz", "dx12", "dy12",
                                                            "dz12")]
    return m, [A.DeviceArray.from_host(a) for a in arrs]


@pytest.mark.parametrize("grid", ["channel", "uniform"])
def test_linebyline_matches_oracle(grid):
    from vtkmakecompare_amd import api as A
    A.init()
    g = G.channel_grid(24, 17, 12) if grid == "channel" else G.set_c(20, 13, 9, 2.0, 1.5, 1.0)
    ut, vt, wt, p = _case(g, 3)
    m, (du, dv, dw, dp) = _dev(A, g, ut, vt, wt, p)
    dt = 1e-3
    want = p.copy()
    for _ in range(3):
        A.linebyline(g.nx, g.ny, g.nz, du, dv, dw, *m, dt, dp)
        oracle().o_linebyline(g.nx, g.ny, g.nz, P(ut), P(vt), P(wt),
                              *[P(getattr(g, k)) for k in ("dx", "dy", "dz", "dx12", "dy12",
                                                          "dz12")], dt, P(want))
    np.testing.assert_array_equal(dp.download(), want)


def test_solve_p_matches_oracle():
    from vtkmakecompare_amd import api as A
    A.init()
    g = G.set_c(12, 10, 8, 2.0, 2.0, 2.0)
    ut, vt, wt, p = _case(g, 5, 1e-3, compatible=True)
    m, (du, dv, dw, dp) = _dev(A, g, ut, vt, wt, p)
    dt = 1e-3
    res, rmax, rint, rvar, it = A.solve_p(dp, du, dv, dw, dt, *m, g.nx, g.ny, g.nz)
    want = p.copy()
    r = [ctypes.c_double(0.0) for _ in range(4)]
    oit = ctypes.c_int(0)
    mets = [P(getattr(g, k)) for k in ("dx", "dy", "dz", "dx12", "dy12", "dz12")]
    oracle().o_solve_p(P(want), P(ut), P(vt), P(wt), dt, *mets, g.nx, g.ny, g.nz,
                       *[ctypes.byref(x) for x in r], 100000, ctypes.byref(oit))
    print(f"solve_p: {it} iterations (oracle {oit.value}), res_max {rmax:.3e}")
    assert it == oit.value and 1 <= it < 100000
    np.testing.assert_array_equal(dp.download(), want)
    assert res == r[0].value and rmax == r[1].value and rmax < 1e-5
    assert rvar == pytest.approx(r[3].value, rel=1e-12)
    # res_int sums signed terms: bound its rounding by the terms' sizes
    assert abs(rint - r[2].value) <= 1e-12 * g.npoints * rmax * float(
        np.max(g.dx) * np.max(g.dy) * np.max(g.dz))
    # a bounded run stops where it is told
    dp2 = A.DeviceArray.from_host(p)
    *_, it2 = A.solve_p(dp2, du, dv, dw, dt, *m, g.nx, g.ny, g.nz, max_iterations=3)
    assert it2 == 3


def test_solve_p_refuses_no_iterations():
    """An iteration limit below 1 is refused (no residual would be set)."""
    from vtkmakecompare_amd import api as A
    A.init()
    g = G.set_c(8, 6, 6, 2.0, 2.0, 2.0)
    ut, vt, wt, p = _case(g, 7)
    m, (du, dv, dw, dp) = _dev(A, g, ut, vt, wt, p)
    with pytest.raises(A.VKError) as e:
        A.solve_p(dp, du, dv, dw, 1e-3, *m, g.nx, g.ny, g.nz, max_iterations=0)
    assert e.value.status == 5  # VK_ERR_ARGUMENT
    np.testing.assert_array_equal(dp.download(), p)
