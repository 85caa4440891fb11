"""The fused steps spread their work over four streams: the temperature's
explicit kernel, sweeps and halo on one, the u and v sweeps on two more,
the rest on the library stream, joined by events.  While kernels are
timed (vk_profile_begin) the same work is issued on the library stream one
kernel after another.  Every kernel is deterministic and the two orders
have the same data dependencies, so the fields must agree bit for bit: a
missing wait between the streams would show here as a difference.
"""
import numpy as np
import pytest

from oracle_bind import oracle_levelset
from vtkmakecompare_amd import grid as G

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "t", "u_temp", "v_temp", "w_temp", "u_ta", "v_ta", "w_ta",
          "u_pre", "v_pre", "w_pre", "t_pre", "divu")


@pytest.fixture(scope="module")
def case():
    from vtkmakecompare_amd import api as A
    A.init()
    cfg = G.small_config(48, 33, 32, seed=0x5EED00D1,
                         spheres=[(2.0, -0.2, 1.0, 0.45), (4.5, 0.4, 2.2, 0.3)])
    g = cfg.grid()
    st = G.initial_state(cfg, g)
    _, cv = oracle_levelset(g, st["phi0"], 200)
    return A, cfg, g, st, cv


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_streams_match_one_stream(case, precision):
    A, cfg, g, st, cv = case
    from vtkmakecompare_amd._lib import check, lib
    nsteps = 6
    a = A.ChannelRun(cfg, g, st, cellvalue=cv, precision=precision)
    a.step(nsteps)
    b = A.ChannelRun(cfg, g, st, cellvalue=cv, precision=precision)
    check(lib().vk_profile_begin(), "profile")
    try:
        b.step(nsteps)
        A.synchronize()
    finally:
        check(lib().vk_profile_end(), "profile")
    for k in FIELDS:
        np.testing.assert_array_equal(a.get(k), b.get(k), err_msg=k)
