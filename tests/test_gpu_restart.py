"""A restart that continues a run bit for bit.  READFIELD's file
(nse_dominik.f:2769-2793) lacks the AB2 history and the step counter, so
a run restarted from it takes an Euler step (l = 1) where the
uninterrupted run takes an AB2 step.  The restart record (vk_write_restart
/ vk_read_restart) carries both: a run saved after 3 steps, restored into
a fresh run and stepped 3 more times equals the uninterrupted 6-step run
bit for bit in every field, in both forms of the step.  The READFIELD
route, for contrast, does not.
"""
import numpy as np
import pytest

from oracle_bind import oracle_levelset
from vtkmakecompare_amd import grid as G

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "t", "u_pre", "v_pre", "w_pre", "t_pre", "u_ta", "v_ta", "w_ta",
          "u_temp", "v_temp", "w_temp", "divu", "cellvalue")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_restart_continues_bit_for_bit(tmp_path, precision):
    from vtkmakecompare_amd import api as A
    A.init()
    cfg = G.small_config(32, 25, 24, seed=0x5EED0401)
    g = cfg.grid()
    st = G.initial_state(cfg, g)
    _, cv = oracle_levelset(g, st["phi0"], 100)
    whole = A.ChannelRun(cfg, g, st, cellvalue=cv, precision=precision)
    whole.step(6)
    first = A.ChannelRun(cfg, g, st, cellvalue=cv, precision=precision)
    first.step(3)
    path = str(tmp_path / "run.restart")
    first.save(path)
    first.free()
    # a fresh run from another state, overwritten by the record
    other = G.initial_state(G.small_config(32, 25, 24, seed=0x5EED0402), g)
    second = A.ChannelRun(cfg, g, other, cellvalue=np.zeros_like(cv), precision=precision)
    second.restore(path)
    assert second.l == 3
    second.step(3)
    for k in FIELDS:
        np.testing.assert_array_equal(second.get(k), whole.get(k), err_msg=k)


def test_readfield_restart_is_not_exact(tmp_path):
    """The reference's own restart route loses the AB2 history: its
    continuation differs from the uninterrupted run."""
    from vtkmakecompare_amd import api as A
    A.init()
    cfg = G.small_config(32, 25, 24, seed=0x5EED0403)
    g = cfg.grid()
    st = G.initial_state(cfg, g)
    _, cv = oracle_levelset(g, st["phi0"], 100)
    whole = A.ChannelRun(cfg, g, st, cellvalue=cv)
    whole.step(6)
    first = A.ChannelRun(cfg, g, st, cellvalue=cv)
    first.step(3)
    # READFIELD's content: u, v, w, p, T, u_ta, v_ta, w_ta, with the
    # history re-initialised from the fields as program main does before
    # its loop (nse_dominik.f:246-257) and the counter back at 1
    st2 = dict(st)
    for k in ("u", "v", "w", "p", "t", "u_ta", "v_ta", "w_ta"):
        st2[k] = first.get(k)
    for a, b in (("u_pre", "u"), ("v_pre", "v"), ("w_pre", "w"), ("t_pre", "t")):
        st2[a] = st2[b].copy()
    cv3 = first.get("cellvalue")
    second = A.ChannelRun(cfg, g, st2, cellvalue=cv3)
    second.step(3)
    d = float(np.max(np.abs(second.get("u") - whole.get("u"))))
    assert d > 1e-10, d


def test_restart_written_behind_queued_steps(tmp_path):
    """A record written from device arrays while steps are still queued
    holds those steps' results: the FP64 step returns with its kernels in
    flight, and the writer's copies wait for the library's streams (a copy
    on the null stream would not, the library's streams being
    non-blocking)."""
    from vtkmakecompare_amd import api as A
    A.init()
    cfg = G.small_config(96, 65, 64, seed=0x5EED0404)
    g = cfg.grid()
    st = G.initial_state(cfg, g)
    _, cv = oracle_levelset(g, st["phi0"], 100)
    run = A.ChannelRun(cfg, g, st, cellvalue=cv)
    run.step(1)
    run.get("u")  # the first step's one-off set-up done, the queue drained
    path = str(tmp_path / "queued.restart")
    run.step(6)   # queued: vk_time_step returns before its kernels finish
    run.save(path)
    host = {k: np.zeros(g.shape) for k in A.RESTART_FIELDS}
    hcv = np.zeros((4, g.nz, g.ny, g.nx))
    assert A.read_restart(path, g, host, hcv) == 7
    for k in A.RESTART_FIELDS:
        np.testing.assert_array_equal(host[k], run.get(k), err_msg=k)
    np.testing.assert_array_equal(hcv, run.get("cellvalue"))
    run.free()
