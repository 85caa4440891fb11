"""The adjoint gradient of the wall-heat-flux objective with respect to the
penalisation field cellvalue(...,1), against central finite differences of
the same GPU forward step, and the objective itself against the oracle's
trajectory.

J = sum_n dt sum_(i,k) dx dz [2 T^n(i,1,k)/dy12(0) + 2 T^n(i,ny,k)/dy12(ny)].
The backward pass is the exact discrete transpose of the FP64 step, so
<grad J, d> and (J(cv + eps d) - J(cv - eps d)) / (2 eps) agree to the
finite-difference error: bar 1e-6 relative at eps = 1e-6 on seeded random
directions d.  Checkpoint intervals that do not divide the step count
exercise the recompute of a partial last segment.
"""
import numpy as np
import pytest

from oracle_bind import OracleRun, oracle_levelset
from vtkmakecompare_amd import grid as G

pytestmark = pytest.mark.gpu

FD_TOL = 1e-6
EPS = 1e-6


@pytest.fixture(scope="module")
def api():
    from vtkmakecompare_amd import api as A
    A.init()
    return A


def _case(n, seed):
    cfg = G.small_config(*n, seed=seed)
    g = cfg.grid()
    st = G.initial_state(cfg, g)
    _, cv = oracle_levelset(g, st["phi0"], 200)
    return cfg, g, st, cv


def _objective_host(g, dt, ts):
    """J from a list of temperature fields (halo layout)."""
    w = dt * 2.0 * g.dx[1:-1][None, :] * g.dz[1:-1][:, None]
    return float(sum(np.sum(w * (t[1:-1, 1, 1:-1] / g.dy12[0] + t[1:-1, g.ny, 1:-1] /
                                 g.dy12[g.ny])) for t in ts))


def test_objective_matches_oracle_trajectory(api):
    cfg, g, st, cv = _case((16, 17, 16), 0x5EED0201)
    nsteps = 5
    prob = api.AdjointProblem(cfg, g, st, cv, nsteps)
    j_gpu = prob.objective()
    st_o = dict(st)
    st_o["cellvalue"] = cv.copy()
    ora = OracleRun(cfg, g, st_o)
    ts = []
    for _ in range(nsteps):
        ora.step(1)
        ts.append(ora["t"].copy())
    j_ora = _objective_host(g, cfg.dt, ts)
    assert abs(j_gpu - j_ora) <= 1e-12 * abs(j_ora), (j_gpu, j_ora)


@pytest.mark.parametrize("n,nsteps,every", [((16, 17, 16), 6, 4), ((24, 13, 12), 5, 2)])
def test_gradient_matches_central_differences(api, n, nsteps, every):
    cfg, g, st, cv = _case(n, 0x5EED0202)
    prob = api.AdjointProblem(cfg, g, st, cv, nsteps)
    j0, grad = prob.gradient(checkpoint_every=every)
    assert abs(j0 - prob.objective()) <= 1e-13 * abs(j0)
    rng = np.random.default_rng(7)
    for q in range(3):
        d = rng.uniform(-1.0, 1.0, size=(g.nz, g.ny, g.nx))
        cp = cv.copy()
        cm = cv.copy()
        cp[0] += EPS * d
        cm[0] += -EPS * d
        fd = (prob.objective(cp) - prob.objective(cm)) / (2.0 * EPS)
        ad = float(np.sum(grad * d))
        print(f"direction {q}: adjoint {ad:.12e} fd {fd:.12e} rel {abs(ad - fd) / abs(fd):.2e}")
        assert abs(ad - fd) <= FD_TOL * abs(fd), (q, ad, fd)


def test_segment_tapes_match_per_step_retaping(api, monkeypatch):
    """The two backward schedules of the checkpointed driver give the same
    gradient bit for bit: every step of a segment taped while the segment
    is recomputed (where the tapes fit, as here), and each step rerun with
    one step's tape (VK_ADJOINT_SEGMENT_TAPES=0).  Seven steps in segments
    of three include a partial last segment."""
    cfg, g, st, cv = _case((16, 17, 16), 0x5EED0203)
    prob = api.AdjointProblem(cfg, g, st, cv, 7)
    j1, g1 = prob.gradient(checkpoint_every=3)
    monkeypatch.setenv("VK_ADJOINT_SEGMENT_TAPES", "0")
    j2, g2 = prob.gradient(checkpoint_every=3)
    assert j1 == j2
    np.testing.assert_array_equal(g1, g2)
