"""The adjoint256 workload at its full size: 256 x 129 x 256 channel, FP64,
4 seeded spheres as the level set, 200 steps, a checkpoint every 20 steps
with each segment recomputed in the backward pass.  The objective is the
wall heat flux integrated over the walls and the run,
J = sum_n dt sum_(i,k) dx dz [2 T^n(i,1,k)/dy12(0) + 2 T^n(i,ny,k)/dy12(ny)],
and the control the penalisation field cellvalue(...,1) (the field
adjoint.f's operators are meant to differentiate through).

The gradient is the exact discrete adjoint of the FP64 step, so on each of
3 seeded random directions d, <grad J, d> agrees with the central
difference (J(cv + eps d) - J(cv - eps d)) / (2 eps) at eps = 1e-6 to the
difference's own error: bar 1e-6 relative.
"""
import numpy as np
import pytest

from vtkmakecompare_amd import grid as G

pytestmark = pytest.mark.gpu

FD_TOL = 1e-6
EPS = 1e-6
EVERY = 20
DIRECTIONS = 3


def test_adjoint256_gradient_matches_central_differences():
    from vtkmakecompare_amd import api as A
    A.init()
    cfg = G.make_config("adjoint256")
    g = cfg.grid()
    assert (g.nx, g.ny, g.nz, cfg.steps) == (256, 129, 256, 200)
    st = G.initial_state(cfg, g)
    cv = A.levelset_cellvalue(g, st["phi0"])
    prob = A.AdjointProblem(cfg, g, st, cv, cfg.steps)
    print(f"adjoint256: gradient, {cfg.steps} steps, checkpoint every {EVERY}", flush=True)
    j0, grad = prob.gradient(checkpoint_every=EVERY)
    assert np.all(np.isfinite(grad))
    j_fwd = prob.objective()
    assert abs(j0 - j_fwd) <= 1e-13 * abs(j_fwd), (j0, j_fwd)
    rng = np.random.default_rng(cfg.seed)
    for q in range(DIRECTIONS):
        d = rng.uniform(-1.0, 1.0, size=(g.nz, g.ny, g.nx))
        cp = cv.copy()
        cm = cv.copy()
        cp[0] += EPS * d
        cm[0] -= EPS * d
        fd = (prob.objective(cp) - prob.objective(cm)) / (2.0 * EPS)
        ad = float(np.sum(grad * d))
        rel = abs(ad - fd) / abs(fd)
        print(f"adjoint256 direction {q}: adjoint {ad:.12e} fd {fd:.12e} rel {rel:.2e}",
              flush=True)
        assert rel <= FD_TOL, (q, ad, fd)
