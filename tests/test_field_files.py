"""OUTPUT, OUTFIELD and READFIELD (nse_dominik.f:2606-2793), the
reference's field files, against the reference's own compiled routines:
the written files byte for byte (temperature_check.dat, list-directed
text, number for number), and the read fields bit for bit.  READFIELD
reads the field.u that OUTPUT writes; OUTFIELD writes a tid record first,
which the library's reader skips, so that file reads back too (the round
trip below).  Host arrays, so no GPU is touched.
"""
import os
import shutil

import numpy as np
import pytest

from oracle_bind import D, I, P, ref
from vtkmakecompare_amd import api as A
from vtkmakecompare_amd import grid as G

NAMES = ("u", "v", "w", "p", "t", "phi", "u_ta", "v_ta", "w_ta")
needs_ref = pytest.mark.skipif(ref() is None, reason="reference library not built")


def state(g, seed):
    rng = np.random.default_rng(seed)
    return {k: rng.normal(size=g.shape) for k in NAMES}


class chdir:
    def __init__(self, d):
        self.d = d

    def __enter__(self):
        self.old = os.getcwd()
        os.chdir(self.d)

    def __exit__(self, *a):
        os.chdir(self.old)


@needs_ref
@pytest.mark.parametrize("n", [(6, 5, 4), (9, 4, 7)])
def test_outfield_bytes_match_reference(tmp_path, n):
    g = G.channel_grid(*n)
    st = state(g, 1)
    tid = 12.375
    for d in ("ours", "ref"):
        (tmp_path / d / "nseoutput2").mkdir(parents=True)
    with chdir(tmp_path / "ours"):
        A.outfield(g, tid, st)
    with chdir(tmp_path / "ref"):
        s = st
        ref().outfield_(P(g.x), P(g.y), P(g.z), P(s["u"]), P(s["v"]), P(s["w"]), P(s["p"]),
                        I(g.nx), I(g.ny), I(g.nz), D(tid), P(s["t"]), P(s["phi"]),
                        P(s["u_ta"]), P(s["v_ta"]), P(s["w_ta"]))
    for f in ("field.u", "nseoutput2/grid.u"):
        assert (tmp_path / "ours" / f).read_bytes() == (tmp_path / "ref" / f).read_bytes(), f


@needs_ref
def test_readfield_matches_reference(tmp_path):
    g = G.channel_grid(7, 5, 6)
    st = state(g, 2)
    (tmp_path / "nseoutput2").mkdir()
    with chdir(tmp_path):
        A.outfield(g, 1.5, st)
        # the file the reference's READFIELD takes: the per-point records only
        raw = (tmp_path / "field.u").read_bytes()
        (tmp_path / "nseoutput2" / "field.u").write_bytes(raw[16:])
        ours = A.readfield(g, {k: np.zeros(g.shape) for k in NAMES})
        theirs = {k: np.zeros(g.shape) for k in NAMES}
        t = theirs
        ref().readfield_(P(t["u"]), P(t["v"]), P(t["w"]), P(t["p"]), I(g.nx), I(g.ny), I(g.nz),
                         P(t["t"]), P(t["phi"]), P(t["u_ta"]), P(t["v_ta"]), P(t["w_ta"]))
    for k in NAMES:
        np.testing.assert_array_equal(ours[k], theirs[k], err_msg=k)
        np.testing.assert_array_equal(ours[k], st[k], err_msg=k)


def test_outfield_readfield_round_trip(tmp_path):
    g = G.channel_grid(8, 6, 5)
    st = state(g, 3)
    (tmp_path / "nseoutput2").mkdir()
    with chdir(tmp_path):
        A.outfield(g, 3.0, st)
        shutil.copy("field.u", "nseoutput2/field.u")  # with its tid record
        back = A.readfield(g, {k: np.zeros(g.shape) for k in NAMES})
        # a truncated file is refused
        raw = open("nseoutput2/field.u", "rb").read()
        open("nseoutput2/field.u", "wb").write(raw[:-40])
        with pytest.raises(A.VKError):
            A.readfield(g, {k: np.zeros(g.shape) for k in NAMES})
    for k in NAMES:
        np.testing.assert_array_equal(back[k], st[k], err_msg=k)


@needs_ref
def test_output_matches_reference(tmp_path):
    g = G.channel_grid(7, 5, 6)
    rng = np.random.default_rng(9)
    names = ("u", "v", "w", "p", "t", "u_ta", "v_ta", "w_ta", "phi", "divu")
    f = {k: rng.normal(size=g.shape) for k in names}
    f["u"][2, 3, 3] = -f["u"][2, 3, 2]  # a face average of exactly 0
    f["v"][4, 1, 5] = 1e-31 - f["v"][4, 0, 5]  # below the 1e-30 cut
    cv = rng.uniform(size=(4, g.nz, g.ny, g.nx))
    for d in ("ours", "ref"):
        (tmp_path / d).mkdir()
    with chdir(tmp_path / "ours"):
        A.output(g, f, cv)
    with chdir(tmp_path / "ref"):
        ref().output_(P(f["u"]), P(f["v"]), P(f["w"]), P(f["p"]), P(g.x), P(g.y), P(g.z),
                      I(g.nx), I(g.ny), I(g.nz), P(f["t"]), P(cv), P(f["u_ta"]), P(f["v_ta"]),
                      P(f["w_ta"]), P(f["phi"]), P(f["divu"]))
    files = sorted(os.listdir(tmp_path / "ref" / "nseoutput2"))
    assert files == sorted(os.listdir(tmp_path / "ours" / "nseoutput2"))
    assert len(files) == 15
    for name in files:
        a = (tmp_path / "ours" / "nseoutput2" / name).read_bytes()
        b = (tmp_path / "ref" / "nseoutput2" / name).read_bytes()
        assert a == b, name
    ta = [float(x) for x in (tmp_path / "ours" / "temperature_check.dat").read_text().split()]
    tb = [float(x) for x in (tmp_path / "ref" / "temperature_check.dat").read_text().split()]
    assert ta == tb and len(ta) == 2 * (g.ny + 2)
    # OUTPUT's field.u is the restart READFIELD reads
    with chdir(tmp_path / "ours"):
        back = A.readfield(g, {k: np.zeros(g.shape) for k in NAMES})
    for k in NAMES:
        np.testing.assert_array_equal(back[k], f[k], err_msg=k)


def test_field_file_wrappers_refuse_wrong_shapes(tmp_path):
    """The wrappers check every array's shape before the library touches
    it: the library moves (nx+2)(ny+2)(nz+2) doubles per field and
    4 nx ny nz of cellvalue, so a short array would be over-read or
    overwritten.  Nothing is written on a refusal."""
    g = G.channel_grid(6, 5, 4)
    good = state(g, 3)
    cv = np.zeros((4, g.nz, g.ny, g.nx))
    with chdir(tmp_path):
        for bad_name in ("u", "w_ta"):
            bad = dict(good)
            bad[bad_name] = np.zeros((g.nz + 2, g.ny + 2, g.nx + 1))
            with pytest.raises(ValueError, match=bad_name):
                A.outfield(g, 1.0, bad)
            with pytest.raises(ValueError, match=bad_name):
                A.readfield(g, bad)
        with pytest.raises(ValueError, match="cellvalue"):
            A.output(g, dict(good, divu=np.zeros(g.shape)), cv[:3])
        with pytest.raises(ValueError, match="float64"):
            A.output(g, dict(good, divu=np.zeros(g.shape, dtype=np.float32)), cv)
        short = G.channel_grid(6, 5, 4)
        short.y = short.y[:-1]
        with pytest.raises(ValueError, match="y"):
            A.outfield(short, 1.0, good)
    assert not any(tmp_path.iterdir())
