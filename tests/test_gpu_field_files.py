"""OUTFIELD and READFIELD on registered (device-resident) host arrays: the
file carries the device copies, brought back first, and a read refreshes
the device copies.  SET_BC on the mirrors changes the halos on the device
only, so a stale host copy would show in the file."""
import ctypes
import os

import numpy as np
import pytest

from oracle_bind import P, oracle
from vtkmakecompare_amd import grid as G

pytestmark = pytest.mark.gpu
NAMES = ("u", "v", "w", "p", "t", "phi", "u_ta", "v_ta", "w_ta")


@pytest.fixture(scope="module")
def api():
    from vtkmakecompare_amd import api as A
    A.init()
    return A


def test_field_files_follow_the_device(api, tmp_path):
    from vtkmakecompare_amd._lib import check, lib
    g = G.channel_grid(12, 9, 8)
    rng = np.random.default_rng(6)
    host = {k: rng.normal(size=g.shape) for k in NAMES}
    want = {k: v.copy() for k, v in host.items()}
    oracle().o_set_bc(P(want["u"]), P(want["v"]), P(want["w"]), P(want["p"]), g.nx, g.ny, g.nz,
                      P(want["t"]))
    nb = ctypes.c_size_t(host["u"].nbytes)
    ptr = {k: host[k].ctypes.data_as(ctypes.c_void_p) for k in NAMES}
    for k in NAMES:
        check(lib().vk_register(ptr[k], ctypes.byref(nb)), "register")
        check(lib().vk_upload(ptr[k]), "upload")
    try:
        I = api.I
        check(lib().vk_set_bc(ptr["u"], ptr["v"], ptr["w"], ptr["p"], I(g.nx), I(g.ny),
                              I(g.nz), ptr["t"]), "set_bc")
        (tmp_path / "nseoutput2").mkdir()
        cwd = os.getcwd()
        os.chdir(tmp_path)
        try:
            api.outfield(g, 2.0, host)
            os.replace("field.u", "nseoutput2/field.u")
            back = api.readfield(g, {k: np.zeros(g.shape) for k in NAMES})
            for k in NAMES:
                np.testing.assert_array_equal(back[k], want[k], err_msg=k)
            # a read into the registered arrays refreshes their device copies
            for k in NAMES:
                host[k][...] = 0.0
            api.readfield(g, host)
            for k in NAMES:
                host[k][...] = -1.0
                check(lib().vk_download(ptr[k]), "download")
                np.testing.assert_array_equal(host[k], want[k], err_msg=k)
        finally:
            os.chdir(cwd)
    finally:
        for k in NAMES:
            lib().vk_unregister(ptr[k])
