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
