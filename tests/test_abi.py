"""The C-ABI library loads and exports every entry point include/fourier_hip.h
declares; parameter validation that needs no device behaves like the
reference's Factory::get.  (CPU only: no compute calls.)"""
import ctypes
import os
import re

import numpy as np
import pytest

from eman2_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "fourier_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fh_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.EXPORTS, "ctypes binding lacks %s" % n


def _create(**kw):
    p = dict(nx=96, ny=96, nz=96, sym=b"c1", mode=b"gauss_2", corners=0, sqrtnorm=0, usessnr=0, verbose=0, device=0)
    p.update(kw)
    prm = _lib.Params(**p)
    h = ctypes.c_void_p()
    rc = _lib.load().fh_create(ctypes.byref(prm), ctypes.byref(h))
    return rc, _lib.load().fh_last_error(None).decode()


@pytest.mark.parametrize("kw,code,word", [
    (dict(nx=97, ny=97, nz=97), 3, "ImageDimension"),
    (dict(nx=96, ny=96, nz=64), 3, "cubic"),
    (dict(nx=34, ny=34, nz=34), 3, "13-smooth"),
    (dict(sym=b"x5"), 2, "symmetry"),
    (dict(sym=b"icos2"), 2, "symmetry"),
    (dict(mode=b"experimental"), 2, "mode"),
    (dict(usessnr=1, corners=1), 1, "usessnr"),
    (dict(nx=24, ny=24, nz=24, mode=b"kaiser_bessel"), 1, "kaiser_bessel"),
])
def test_create_validation(kw, code, word):
    rc, msg = _create(**kw)
    assert rc == code
    assert word in msg


def test_null_arguments():
    L = _lib.load()
    assert L.fh_create(None, None) == 4
    assert L.fh_setup(None) == 4
    assert L.fh_insert_batch(None, None, 0, 0, None, None, None, None, 0) == 4
    assert L.fh_destroy(None) == 0


def test_python_factory_validation():
    from eman2_amd.reconstructor import Reconstructors, FourierHipReconstructor
    assert "fourier_hip" in Reconstructors.get_list()
    with pytest.raises(RuntimeError, match="NotExistingObject"):
        Reconstructors.get("fourier_nope", {"size": (96, 96, 96)})
    with pytest.raises(RuntimeError, match="InvalidParameter"):
        Reconstructors.get("fourier_hip", {"size": (96, 96, 96), "bogus": 1})
    with pytest.raises(RuntimeError, match="InvalidParameter"):
        FourierHipReconstructor({"subvolume": (0, 0, 0, 10, 10, 10)})
    with pytest.raises(RuntimeError, match="ImageDimension"):
        Reconstructors.get("fourier_hip", {"size": (95, 95, 95)})


def test_footprint_grows_with_size():
    L = _lib.load()
    a, b = L.fh_footprint_bytes(96, 1, 256), L.fh_footprint_bytes(384, 1, 256)
    assert 0 < a < b
    assert b >= 12 * 384 * 384 * 193


SYM_NAMES = ["c1", "C5", "d2", "d7", "h5:1:30:2", "h4,1,12.5,3", "tet", "oct", "icos", "icos5"]


@pytest.mark.parametrize("sym", SYM_NAMES)
def test_symmetry_table_matches_oracle(sym):
    """fh_get_symmetries is the oracle's Symmetry3D::get_symmetries bit for bit."""
    from eman2_amd.transform import get_symmetries
    from oracle import oracle as O
    ops = get_symmetries(sym)
    ref = O.sym_get(sym)[:, :, :3]
    assert ops.shape == ref.shape
    np.testing.assert_array_equal(ops, ref)


def test_symmetry_unknown_and_small_buffer():
    from eman2_amd.transform import get_symmetries
    L = _lib.load()
    for s in (b"x3", b"ico", b"h3:2:10:1", b""):
        assert L.fh_get_symmetries(s, None, 0) == -2
    buf = (ctypes.c_float * 9)()
    assert L.fh_get_symmetries(b"c4", ctypes.cast(buf, ctypes.c_void_p), 1) == -1
    with pytest.raises(_lib.FourierHipError, match="NotExistingObjectException"):
        get_symmetries("ico")
