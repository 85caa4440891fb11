import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


# Guard zones around every device allocation of the library (vk_guard_check
# in include/vtkgpu.h): on for the whole session, read by vk_init, and
# checked after every GPU test, so a kernel that writes past either end of a
# buffer fails the test that ran it.
os.environ.setdefault("VK_GUARD", "1")


@pytest.fixture(autouse=True)
def _guard_zones(request):
    yield
    if request.node.get_closest_marker("gpu") is None:
        return
    from vtkmakecompare_amd import _lib
    if _lib._lib is None:
        return
    import ctypes
    n = ctypes.c_int(0)
    rc = _lib._lib.vk_guard_check(ctypes.byref(n))
    if rc in (0, 1):  # clean, or the library is not initialised
        return
    msg = _lib._lib.vk_last_error().decode(errors="replace")
    pytest.fail(f"after {request.node.nodeid}: status {rc}, {n.value} zone(s): {msg}")
