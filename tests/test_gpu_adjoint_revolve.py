"""The adjoint gradient by binomial checkpointing (vk_adjoint_gradient_revolve).

- FP64 trajectory: the objective and the gradient equal the two-level
  checkpoint driver's (vk_adjoint_gradient) bit for bit, for every number
  of snapshots from none (every state recomputed from the start) to
  nsteps - 1 (every state kept): the schedule hands each backward step
  exactly the state before it.
- the schedule's cost at its two closed-form ends: no snapshot recomputes
  n (n - 1) / 2 steps, n - 1 snapshots recompute n - 1; and the steps the
  call ran are the objective pass plus the schedule's planned steps.
- mixed-precision trajectory: J against the FP64 trajectory's J, and the
  gradient (the FP64 adjoint linearised about the mixed trajectory)
  against the FP64 adjoint: J within 1e-6 and the gradient within 1e-5
  normwise (measured 2e-8 and 1.3e-7 at 32x25x24 over 6 steps), the order
  of the mixed form's own state error against FP64.
"""
import numpy as np
import pytest

from oracle_bind import oracle_levelset
from vtkmakecompare_amd import grid as G

pytestmark = pytest.mark.gpu


def _problem(A, n, nsteps, seed):
    cfg = G.small_config(*n, seed=seed)
    g = cfg.grid()
    st = G.initial_state(cfg, g)
    _, cv = oracle_levelset(g, st["phi0"], 200)
    return g, A.AdjointProblem(cfg, g, st, cv, nsteps)


def test_revolve_matches_two_level_checkpoints():
    from vtkmakecompare_amd import api as A
    A.init()
    N = 7
    g, prob = _problem(A, (16, 17, 16), N, 0x5EED0501)
    j2, grad2 = prob.gradient(checkpoint_every=3)
    for s in (0, 1, 2, 3, N - 1):
        j, grad, st = prob.gradient_revolve(snapshots=s)
        print(f"snapshots {s}: {st}")
        assert j == j2, (s, j, j2)
        np.testing.assert_array_equal(grad, grad2, err_msg=f"snapshots {s}")
        assert st["snapshots"] == s
        assert st["forward_steps"] == N + st["recomputed_steps"]
        if s == 0:
            assert st["recomputed_steps"] == N * (N - 1) // 2
        if s == N - 1:
            assert st["recomputed_steps"] == N - 1
    # snapshots = -1: as many as fit, capped at nsteps - 1
    j, grad, st = prob.gradient_revolve(snapshots=-1)
    assert st["snapshots"] == N - 1 and j == j2
    np.testing.assert_array_equal(grad, grad2)


def test_revolve_mixed_trajectory_against_fp64():
    from vtkmakecompare_amd import api as A
    A.init()
    N = 6
    g, prob = _problem(A, (32, 25, 24), N, 0x5EED0502)
    j64, g64, _ = prob.gradient_revolve(snapshots=2)
    jmx, gmx, st = prob.gradient_revolve(snapshots=2, mixed=True)
    ej = abs(jmx - j64) / abs(j64)
    eg = float(np.max(np.abs(gmx - g64)) / np.max(np.abs(g64)))
    print(f"mixed trajectory: J rel err {ej:.2e}, gradient normwise rel err {eg:.2e}, {st}")
    assert ej < 1e-6, ej
    assert eg < 1e-5, eg
