"""EMAN2 HDF5 image files (libEM/io/hdfio2.cpp layout) through the HDF5 C
library.  The reference ships no HDF5 files: parity unpinned against
EMAN2-written files; pinned here by round trips, the layout the reader
expects, and the closed forms of the bit reduction (renderer.h:74-110,
hdfio2.cpp:1251-1252).  (CPU only.)"""
import numpy as np
import pytest

from eman2_amd import hdf

try:
    hdf._lib()
    HAVE = True
except OSError:
    HAVE = False
pytestmark = pytest.mark.skipif(not HAVE, reason="libhdf5 not present")


def test_stack_round_trip_with_headers(tmp_path):
    rng = np.random.default_rng(1)
    imgs = rng.standard_normal((5, 12, 10)).astype(np.float32)
    xf = np.arange(12, dtype=np.float32).reshape(3, 4) / 7
    heads = [{"apix_x": 1.5, "ctf_defocus": np.float64(1.25 + i), "class_id": i, "flip": bool(i % 2),
              "source": "ptcl_%d" % i, "xform.projection": xf * (i + 1), "ssnr": np.linspace(0, 1, 4, dtype=np.float32),
              "idx": np.arange(3)} for i in range(5)]
    p = str(tmp_path / "s.hdf")
    hdf.write_images(p, imgs, heads)
    s = hdf.HdfStack(p)
    assert len(s) == 5
    for i in range(5):
        np.testing.assert_array_equal(s.read_image(i), imgs[i])
        h = s.header(i)
        assert h["nx"] == 10 and h["ny"] == 12 and h["nz"] == 1
        assert h["apix_x"] == np.float32(1.5) and h["ctf_defocus"] == 1.25 + i
        assert h["class_id"] == i and h["flip"] == bool(i % 2) and h["source"] == "ptcl_%d" % i
        np.testing.assert_array_equal(h["xform.projection"], xf * (i + 1))
        np.testing.assert_array_equal(h["ssnr"], np.linspace(0, 1, 4, dtype=np.float32))
        np.testing.assert_array_equal(h["idx"], np.arange(3))
    s.close()
    a, hs = hdf.read_stack(p)
    np.testing.assert_array_equal(a, imgs)


def test_volume_float_deflate_lossless(tmp_path):
    v = np.random.default_rng(2).standard_normal((9, 8, 10)).astype(np.float32)
    p = str(tmp_path / "v.hdf")
    hdf.write_compressed(p, v, bits=0)
    s = hdf.HdfStack(p)
    np.testing.assert_array_equal(s.read_image(0), v)
    assert "stored_renderbits" not in s.header(0)


@pytest.mark.parametrize("bits", [4, 8, 10, 16])
def test_bit_reduction_closed_form(tmp_path, bits):
    """each stored value is round((v - rmin) / (rmax - rmin) (2^bits - 1)),
    read back as (i / (2^bits - 1)) (rmax - rmin) + rmin: within half a step
    of v inside the range, the range ends outside it"""
    rng = np.random.default_rng(bits)
    v = rng.standard_normal((6, 7, 8)).astype(np.float32)
    v[0, 0, 0], v[0, 0, 1] = 40.0, -40.0         # outliers
    p = str(tmp_path / ("c%d.hdf" % bits))
    hdf.write_compressed(p, v, bits=bits, nooutliers=True)
    s = hdf.HdfStack(p)
    h = s.header(0)
    rmin, rmax = h["stored_rendermin"], h["stored_rendermax"]
    assert h["stored_renderbits"] == bits and rmin < 0 < rmax and h["stored_truncated"] >= 2
    got = s.read_image(0)
    step = (rmax - rmin) / ((1 << bits) - 1)
    inside = (v >= rmin) & (v <= rmax)
    assert np.abs(got - v)[inside].max() <= 0.5 * step * (1 + 1e-5) + 1e-6
    np.testing.assert_allclose(got[v > rmax], rmax, rtol=1e-6)
    np.testing.assert_allclose(got[v < rmin], rmin, rtol=1e-6)
    ints, trunc = hdf.render(v, rmin, rmax, bits)
    np.testing.assert_array_equal(got, hdf.unrender(ints, rmin, rmax, bits))
    assert trunc == h["stored_truncated"]


def test_render_limits_keep_zero_and_integers():
    """getRenderMinMax: with min < 0 < max the range is an integral number of
    steps from zero (a masked map's zeros stay exactly 0); integer images
    narrower than 2^bits are stored exactly"""
    v = np.concatenate([np.linspace(-1.3, 2.9, 101), np.zeros(50)]).astype(np.float32)
    rmin, rmax, b = hdf.render_limits(v, 8)
    ints, _ = hdf.render(v, rmin, rmax, b)
    back = hdf.unrender(ints, rmin, rmax, b)
    assert np.all(np.abs(back[v == 0]) < 1e-6 * (rmax - rmin))
    k = np.arange(-20, 180, dtype=np.float32)
    rmin, rmax, b = hdf.render_limits(k, 8)
    assert (rmin, rmax) == (-20.0, 235.0)
    ints, _ = hdf.render(k, rmin, rmax, b)
    np.testing.assert_array_equal(hdf.unrender(ints, rmin, rmax, b), k)


def test_outlier_limits_histogram_walk():
    """the nooutliers range: the first histogram bin from the bottom whose
    cumulative count passes max(n/20000, 8), and from the top likewise but
    never below bin 2049 (the walks stop at the middle of the histogram)"""
    v = np.random.default_rng(3).standard_normal(200000).astype(np.float32)
    lo, hi = hdf.outlier_limits(v)
    h0, h1 = float(v.min()), float(v.max())
    hs = (h1 - h0) / 4095
    assert np.count_nonzero(v < lo - hs) <= 10 and np.count_nonzero(v > hi + hs) <= 10
    assert np.count_nonzero(v <= lo + hs) > 10 and np.count_nonzero(v >= hi - hs) > 10
    v[:3] = 1e3                                      # far outliers squeeze the histogram
    lo, hi = hdf.outlier_limits(v)
    hs = (1e3 - float(v.min())) / 4095
    assert abs(hi - (2049 * hs + float(v.min()))) < 1e-6 * 1e3


class _Recorder:
    def setup(self):
        pass

    def insert_batch(self, imgs, xforms, weights, defocus=None, ctf=None):
        self.args = (np.array(imgs), np.array(xforms), np.array(weights))

    def finish(self, doift=True):
        return np.full((8, 8, 8), 0.5, np.float32)


def test_reconstruct_from_hdf_stack_and_lst(tmp_path, monkeypatch):
    from eman2_amd import lst
    from eman2_amd.make3d import reconstruct_hdf, reconstruct_lst
    rng = np.random.default_rng(9)
    imgs = rng.standard_normal((4, 8, 8)).astype(np.float32)
    mats = [np.eye(3, 4, dtype=np.float32) * (i + 1) for i in range(4)]
    p = str(tmp_path / "ptcls.hdf")
    hdf.write_images(p, imgs, [{"xform.projection": m, "apix_x": 2.5} for m in mats])
    r = _Recorder()
    out, info = reconstruct_hdf(r, p, out_path=str(tmp_path / "map.hdf"), niter=1)
    np.testing.assert_array_equal(r.args[0], imgs)
    np.testing.assert_array_equal(r.args[1], np.stack(mats))
    s = hdf.HdfStack(str(tmp_path / "map.hdf"))
    h = s.header(0)
    assert h["apix_x"] == np.float32(2.5) and h["stored_renderbits"] == 10
    np.testing.assert_allclose(s.read_image(0), 0.5, atol=1e-3)
    # an #LSX list naming the HDF stack, relative to the working directory
    monkeypatch.chdir(tmp_path)
    lst.write("sel.lst", [(3, "ptcls.hdf", {"xform.projection": mats[0]}), (1, "ptcls.hdf", {"xform.projection": mats[2]})])
    r2 = _Recorder()
    reconstruct_lst(r2, "sel.lst", niter=1)
    np.testing.assert_array_equal(r2.args[0], imgs[[3, 1]])


def test_render_rounds_half_away_from_zero():
    """roundf, not round-half-even: 2.5 steps -> 3, 1.5 -> 2"""
    ints, _ = hdf.render(np.array([2.5, 1.5, 0.5], np.float32), 0.0, 255.0, 8)
    assert ints.tolist() == [3, 2, 1]
