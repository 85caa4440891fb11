"""The channel64 configuration (64 x 65 x 64, one sphere of radius 0.25,
Re 2800, dt 1e-3) run for its full 100 steps: the FP64 step against the
oracle at 1e-12 and the mixed step at its velocity bar, field by field
(normwise relative error over the interior, max|a - b| / max|b|).

Both runs start from the same cellvalue (set_levelset on the device,
whose own parity is a test of its own), so the comparison covers the
time loop alone.
"""
import numpy as np
import pytest

from oracle_bind import OracleRun
from vtkmakecompare_amd import grid as G

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "t", "u_ta", "v_ta", "w_ta")


def nrel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


@pytest.fixture(scope="module")
def runs():
    from vtkmakecompare_amd import api as A
    A.init()
    cfg = G.make_config("channel64")
    g = cfg.grid()
    st = G.initial_state(cfg, g)
    cv = A.levelset_cellvalue(g, st["phi0"])
    st_o = dict(st)
    st_o["cellvalue"] = cv.copy()
    ora = OracleRun(cfg, g, st_o)
    ora.step(cfg.steps)
    out = {}
    for precision in ("fp64", "mixed"):
        run = A.ChannelRun(cfg, g, st, cellvalue=cv, precision=precision)
        run.step(cfg.steps)
        out[precision] = {k: run.get(k) for k in FIELDS + ("divu",)}
    return cfg, g, ora, out


def test_channel64_fp64_full_run(runs):
    cfg, g, ora, out = runs
    errs = {k: nrel(interior(out["fp64"][k]), interior(ora[k])) for k in FIELDS}
    print("channel64 fp64 after", cfg.steps, "steps:", errs)
    for k, e in errs.items():
        assert e < 1e-12, (k, e)


def test_channel64_mixed_full_run(runs):
    cfg, g, ora, out = runs
    errs = {k: nrel(interior(out["mixed"][k]), interior(ora[k])) for k in FIELDS}
    print("channel64 mixed after", cfg.steps, "steps:", errs)
    for k in ("u", "v", "w"):
        assert errs[k] < 1e-6, (k, errs[k])
    assert errs["t"] < 1e-5
    assert errs["p"] < 1e-5


def test_channel64_divergence(runs):
    """The projected velocity is discretely divergence-free in both forms."""
    cfg, g, ora, out = runs
    umax = np.max(np.abs(interior(ora["u"])))
    for precision in ("fp64", "mixed"):
        div = np.max(np.abs(interior(out[precision]["divu"])))
        assert div < 1e-10 * umax / g.dy[1], (precision, div)
