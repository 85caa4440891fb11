"""Failure handling at the C ABI.

The reference has no error channel: a blown-up run shows as NaNs in its
output files.  The library returns a status from every entry point and
refuses, before anything is launched, the inputs its kernels cannot
take.  Each case here is one of those refusals or the NaN/Inf check of
vk_field_absmax; no kernel is ever handed an invalid input.
"""
import ctypes

import numpy as np
import pytest

from vtkmakecompare_amd import grid as G

pytestmark = pytest.mark.gpu

VK_ERR_NOT_RESIDENT = 2
VK_ERR_BAD_SIZE = 3
VK_ERR_ARGUMENT = 5
VK_ERR_NONFINITE = 6
VK_ERR_GUARD = 7


@pytest.fixture(scope="module")
def api():
    from vtkmakecompare_amd import api as A
    A.init()
    return A


def dev(a):
    from vtkmakecompare_amd.device import DeviceArray
    return DeviceArray.from_host(np.ascontiguousarray(a))


def test_absmax_matches_numpy(api):
    g = G.channel_grid(24, 17, 12)
    rng = np.random.default_rng(3)
    f = rng.normal(size=g.shape)
    f[0, :, :] = 1e3  # halo values are not part of the check
    d = dev(f)
    got = api.field_absmax(d, g.nx, g.ny, g.nz)
    assert got == np.max(np.abs(f[1:-1, 1:-1, 1:-1]))


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
@pytest.mark.parametrize("where", [(1, 1, 1), (12, 17, 24), (5, 9, 3)])
def test_absmax_reports_nonfinite(api, bad, where):
    g = G.channel_grid(24, 17, 12)
    f = np.ones(g.shape)
    f[where] = bad
    with pytest.raises(api.VKError) as e:
        api.field_absmax(dev(f), g.nx, g.ny, g.nz)
    assert e.value.status == VK_ERR_NONFINITE


def test_absmax_ignores_halo_nan(api):
    g = G.channel_grid(24, 17, 12)
    f = np.full(g.shape, 0.5)
    f[0, 0, 0] = np.nan
    f[-1, 3, 4] = np.inf
    assert api.field_absmax(dev(f), g.nx, g.ny, g.nz) == 0.5


def _step_args(api, g, d, cfg, nx=None):
    nx = g.nx if nx is None else nx
    I, D = api.I, api.D
    return (d["u"], d["v"], d["w"], d["p"], d["t"], d["u_pre"], d["v_pre"], d["w_pre"],
            d["t_pre"], d["u_ta"], d["v_ta"], d["w_ta"], d["u_temp"], d["v_temp"],
            d["w_temp"], d["divu"], d["cellvalue"], d["rhocp"], d["thc"], None, None, None,
            d["dx"], d["dy"], d["dz"], d["dx12"], d["dy12"], d["dz12"], I(nx), I(g.ny),
            I(g.nz), D(cfg.dt), D(cfg.re), D(cfg.sc), D(cfg.eta), D(g.xl), D(g.zl), I(1))


def _resident_case(api, n=(16, 17, 16)):
    cfg = G.small_config(*n)
    g = cfg.grid()
    st = G.initial_state(cfg, g)
    d = {k: dev(st[k]) for k in api.STEP_FIELDS if not k.endswith("_ep2")}
    d["cellvalue"] = dev(np.zeros((4, g.nz, g.ny, g.nx)))
    for k in ("dx", "dy", "dz", "dx12", "dy12", "dz12"):
        d[k] = dev(getattr(g, k))
    return cfg, g, st, d


def test_extent_of_device_arrays(api):
    g = G.channel_grid(16, 17, 16)
    d = dev(np.zeros(g.shape))
    assert api.extent(d) == d.nbytes


def test_time_step_refuses_small_array(api):
    """A device array too small for the grid is refused before launch."""
    cfg, g, st, d = _resident_case(api)
    small = dev(np.zeros(g.shape)[: g.nz // 2])
    # the refusal rests on the runtime's allocation range: it must report
    # less than a field here, or the step below would read past the array
    assert api.extent(small) < np.zeros(g.shape).nbytes
    d["w_pre"] = small
    with pytest.raises(api.VKError) as e:
        api.check(api.lib().vk_time_step(*_step_args(api, g, d, cfg)), "time step")
    assert e.value.status == VK_ERR_ARGUMENT
    assert "w_pre" in str(e.value)


def test_time_step_refuses_host_array(api):
    """A host array that was never registered is refused before launch."""
    cfg, g, st, d = _resident_case(api)
    host = np.zeros(g.shape)
    d["u"] = host.ctypes.data_as(ctypes.c_void_p)
    with pytest.raises(api.VKError) as e:
        api.check(api.lib().vk_time_step(*_step_args(api, g, d, cfg)), "time step")
    assert e.value.status == VK_ERR_NOT_RESIDENT


def test_time_step_refuses_tiny_grid(api):
    cfg, g, st, d = _resident_case(api)
    with pytest.raises(api.VKError) as e:
        api.check(api.lib().vk_time_step(*_step_args(api, g, d, cfg, nx=2)), "time step")
    assert e.value.status == VK_ERR_BAD_SIZE


@pytest.mark.parametrize("nx,nz", [(14, 16), (16, 10), (15, 16)])
def test_mkp_refuses_unsupported_fft_lengths(api, nx, nz):
    """The FFT takes nx even with nx/2 and nz of the form 2^a 3^b."""
    g = G.channel_grid(nx, 9, nz)
    f = [dev(np.zeros(g.shape)) for _ in range(14)]
    m = [dev(getattr(g, k)) for k in ("dx", "dy", "dz", "dx12", "dy12", "dz12")]
    I, D = api.I, api.D
    with pytest.raises(api.VKError) as e:
        api.mkp(f[0], f[1], f[2], f[3], 1e-3, m[0], m[1], m[2], m[3], m[4], m[5], 2800.0,
                g.nx, g.ny, g.nz, 1, f[4], f[5], f[6], f[7], f[8], f[9], f[10], f[11], f[12],
                f[13], g.xl, g.zl)
    assert e.value.status == VK_ERR_BAD_SIZE


def test_guard_zones_report_writes_outside_an_allocation(api):
    """vk_guard_check names the allocation and the offsets of bytes written
    in the zones around it.  The bytes are written from the host (vk_memset
    on the zone's own memory, which the guarded allocation holds), so no
    kernel goes out of bounds; a second check finds the zones refilled."""
    from vtkmakecompare_amd._lib import lib
    from vtkmakecompare_amd.device import DeviceArray
    L = lib()
    n = ctypes.c_int(-2)
    assert L.vk_guard_check(ctypes.byref(n)) == 0
    if n.value == -1:  # no zones: then the writes below would leave the allocation
        pytest.skip("guard zones are off (VK_GUARD)")
    assert n.value == 0
    a = DeviceArray((125,))  # 1000 bytes
    ext = ctypes.c_size_t(0)
    assert L.vk_extent(a.ptr, ctypes.byref(ext)) == 0 and ext.value == 1000
    base = a.ptr.value
    zero = ctypes.byref(ctypes.c_int(0))
    assert L.vk_memset(ctypes.c_void_p(base + 1000), zero, ctypes.byref(ctypes.c_size_t(8))) == 0
    assert L.vk_memset(ctypes.c_void_p(base - 16), zero, ctypes.byref(ctypes.c_size_t(4))) == 0
    rc = L.vk_guard_check(ctypes.byref(n))
    msg = L.vk_last_error().decode()
    assert rc == VK_ERR_GUARD and n.value == 2, msg
    assert "vk_malloc (1000 bytes): 8 bytes written after its end, offsets 1000..1007" in msg
    assert "vk_malloc (1000 bytes): 4 bytes written before its start, offsets -16..-13" in msg
    assert L.vk_guard_check(ctypes.byref(n)) == 0 and n.value == 0
    a.free()
