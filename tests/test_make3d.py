"""Host logic of the e2make3d.py reconstruct() loop (programs/e2make3d.py:494-539):
relative qualities and the quality cutoff, on hand-computed cases."""
import numpy as np
import pytest

from eman2_amd.make3d import quality_cutoff, relative_quality


def test_relative_quality_small_window():
    # all five within one +-10 window: each absqual over the mean of all five (0.56);
    # the weight-0 particle gets -1 and stays out of the list
    w = np.array([1, 1, 0, 2, 1.0])
    a = np.array([0.5, 0.7, 0.1, 0.9, 0.6])
    q, ql = relative_quality(w, a)
    np.testing.assert_allclose(q, [0.5 / 0.56, 0.7 / 0.56, -1, 0.9 / 0.56, 0.6 / 0.56])
    assert sorted(ql) == pytest.approx(sorted([0.5 / 0.56, 0.7 / 0.56, 0.9 / 0.56, 0.6 / 0.56]))
    # keep 0.8 of 4: the 3rd largest relative quality
    assert quality_cutoff(a, ql, 0.8, False, False) == pytest.approx(0.6 / 0.56)


def test_relative_quality_running_window():
    # 30 particles of equal weight sorted by absqual: particle k's window is
    # sorted positions [k-10, k+10)
    n = 30
    a = np.linspace(1.0, 3.9, n)
    w = np.ones(n)
    q, _ = relative_quality(w, a)
    for k in (0, 5, 15, 29):
        win = a[max(0, k - 10):min(k + 10, n)]
        assert q[k] == pytest.approx(a[k] / win.mean())


def test_cutoffs_sigma_and_abs():
    a = np.array([1.0, 2.0, 3.0, 4.0])
    ql = [0.5, 1.0, 1.5, 2.0]
    m, s = np.mean(ql), np.std(ql)
    assert quality_cutoff(a, ql, 1.0, True, False) == pytest.approx(m - s)
    assert quality_cutoff(a, ql, 0.5, False, True) == pytest.approx(3.0)     # sorted abs: [-int(0.5*4)] = [-2]
    assert quality_cutoff(a, ql, 0.0, True, 2.0) == pytest.approx(a.mean() - 2.0 * a.std())


def test_threshold_compress_closed_form():
    """CollapseProcessor at the map's mean v with its n-1 standard deviation r:
    x - r above v + r, x + r below v - r, v in between"""
    from eman2_amd.make3d import threshold_compress
    a = np.array([[[0.0, 1.0], [2.0, 10.0]], [[-6.0, 3.0], [4.0, 2.0]]], np.float32)
    v = np.float32(a.astype(np.float64).mean())          # stored as float, as EMData's attributes
    r = np.float32(a.astype(np.float64).std(ddof=1))
    out = threshold_compress(a)
    want = np.where(a > v + r, a - r, np.where(a < v - r, a + r, v))
    np.testing.assert_array_equal(out, want)
    assert out[0, 1, 1] == np.float32(10.0) - r and out[1, 0, 0] == np.float32(-6.0) + r
    assert out[0, 0, 1] == v and r > 4.4


def test_reconstruct_iterative_seeds_each_pass_with_the_compressed_map():
    from eman2_amd.make3d import reconstruct_iterative, threshold_compress

    class Rec:
        def __init__(self):
            self.calls = []
            self.k = 0

        def setup(self):
            self.calls.append(("setup",))

        def setup_seed(self, seed, w):
            self.calls.append(("seed", np.array(seed), w))

        def insert_batch(self, imgs, xf, w, defocus=None, ctf=None):
            self.calls.append(("insert", len(imgs)))

        def finish(self, doift=True):
            self.k += 1
            return np.full((4, 4, 4), float(self.k), np.float32) + np.arange(64, dtype=np.float32).reshape(4, 4, 4)

    r = Rec()
    out = reconstruct_iterative(r, np.zeros((5, 6, 6), np.float32), np.zeros((5, 3, 4), np.float32), niter=3)
    kinds = [c[0] for c in r.calls]
    assert kinds == ["setup", "insert", "seed", "insert", "seed", "insert"]
    first = np.full((4, 4, 4), 1.0, np.float32) + np.arange(64, dtype=np.float32).reshape(4, 4, 4)
    np.testing.assert_array_equal(r.calls[2][1], threshold_compress(first))
    assert r.calls[2][2] == 1.0
    assert out[0, 0, 0] == 3.0     # the last pass's map, not compressed
