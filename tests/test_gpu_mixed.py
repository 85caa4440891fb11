"""The mixed-precision time step against the FP64 oracle.

The mixed form keeps the velocity, u_temp, p and divu in FP64 and the
temperature, AB2 history, u_ta, cellvalue, rhoCp and thc in float; the
explicit phase and the FFTs compute in FP32.  Bars (normwise relative,
max|mixed - oracle| / max|oracle| over the interior):
  velocity u, v, w       <= 1e-6
  pressure p             <= 1e-5
  temperature T          <= 1e-5
and the post-projection divergence at FP64 round-off (<= 1e-10 in
absolute terms on these grids), which is what the FP64 refinement pass of
the projection buys.
"""
import numpy as np
import pytest

from oracle_bind import OracleRun, oracle_levelset
from vtkmakecompare_amd import grid as G

pytestmark = pytest.mark.gpu

VEL_TOL = 1e-6
P_TOL = 1e-5
T_TOL = 1e-5
DIV_TOL = 1e-10


@pytest.fixture(scope="module")
def api():
    from vtkmakecompare_amd import api as A
    A.init()
    return A


def nrel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


def _case(n, seed):
    cfg = G.small_config(*n, seed=seed)
    g = cfg.grid()
    st = G.initial_state(cfg, g)
    _, cv = oracle_levelset(g, st["phi0"], 200)
    return cfg, g, st, cv


@pytest.mark.parametrize("n,nsteps", [((24, 33, 24), 10), ((48, 25, 36), 6)])
def test_mixed_step_against_fp64_oracle(api, n, nsteps):
    cfg, g, st, cv = _case(n, 0x5EED0101)
    run = api.ChannelRun(cfg, g, st, cellvalue=cv, precision="mixed")
    st_o = dict(st)
    st_o["cellvalue"] = cv.copy()
    ora = OracleRun(cfg, g, st_o)
    run.step(nsteps)
    ora.step(nsteps)
    errs = {k: nrel(interior(run.get(k)), interior(ora[k])) for k in ("u", "v", "w", "p", "t")}
    print(f"mixed vs oracle after {nsteps} steps at {n}: " +
          ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    for k in ("u", "v", "w"):
        assert errs[k] < VEL_TOL, (k, errs[k])
    assert errs["p"] < P_TOL, errs["p"]
    assert errs["t"] < T_TOL, errs["t"]
    div = float(np.max(np.abs(interior(run.get("divu")))))
    assert div < DIV_TOL, div


def test_mixed_step_against_fp64_step(api):
    """Same initial state, both GPU forms, a larger grid with the
    stretched y and the radix-3 FFT length."""
    cfg, g, st, cv = _case((96, 49, 48), 0x5EED0102)
    a = api.ChannelRun(cfg, g, st, cellvalue=cv, precision="mixed")
    b = api.ChannelRun(cfg, g, st, cellvalue=cv)
    a.step(8)
    b.step(8)
    for k in ("u", "v", "w"):
        assert nrel(interior(a.get(k)), interior(b.get(k))) < VEL_TOL, k
    assert nrel(interior(a.get("t")), interior(b.get("t"))) < T_TOL
    assert float(np.max(np.abs(interior(a.get("divu"))))) < DIV_TOL
