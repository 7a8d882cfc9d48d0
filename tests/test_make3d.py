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


def test_get_rotation_recovers_the_matrix():
    """Transform.get_rotation("eman") (transform.cpp:852-898): the angles it
    returns rebuild the same rotation, with and without an x mirror or a
    scale, and at alt 0 / 180 (the whole in-plane angle goes to phi)"""
    from eman2_amd.transform import Transform, euler_eman_matrix
    rng = np.random.default_rng(3)
    cases = [tuple(rng.random(3) * (360, 180, 360)) for _ in range(20)] + [(35.0, 0.0, 20.0), (35.0, 180.0, 20.0)]
    for i, (az, alt, phi) in enumerate(cases):
        mirror, scale = i % 3 == 1, 1.0 if i % 4 else 1.25
        t = Transform({"type": "eman", "az": az, "alt": alt, "phi": phi, "mirror": mirror, "scale": scale})
        s, m = t.get_scale_and_mirror()
        assert m == mirror and abs(s - scale) < 1e-6
        r = t.get_rotation("eman")
        back = euler_eman_matrix(r["az"], r["alt"], r["phi"])[:, :3] * np.float32(scale)
        if mirror:
            back[0] *= -1
        np.testing.assert_allclose(back, t.matrix[:, :3], atol=2e-6)
    r0 = Transform({"type": "eman", "az": 35.0, "alt": 0.0, "phi": 20.0}).get_rotation()
    assert r0["az"] == 0.0 and r0["alt"] == 0.0 and abs(r0["phi"] - 55.0) < 1e-4


def test_fill_orientations_grid_and_weights():
    """e2make3dpar.py --fillangle (:666-677, 737-747): step atan2(1, P/2) deg,
    den = floor(fillangle / step) capped at 9, no filling for den <= 1; a den x
    den grid of (alt, az) offsets with factor exp(-(da^2 + dz^2) / (fa/4)^2),
    phi 0, the particle's own translation kept for preprocess_slice"""
    import math

    from eman2_amd.make3d import fill_orientations
    from eman2_amd.transform import Transform, euler_eman_matrix
    m = euler_eman_matrix(30.0, 40.0, 50.0)
    m[:, 3] = (2.5, -1.5, 0.0)
    step = math.degrees(math.atan2(1.0, 192.0))               # P = 384
    assert len(fill_orientations(m, 1.9 * step, 384)) == 1   # den 1: no filling
    np.testing.assert_array_equal(fill_orientations(m, 1.9 * step, 384)[0][0], m)
    assert len(fill_orientations(m, 100.0, 384)) == 81       # den capped at 9
    f = fill_orientations(m, 1.0, 384)                        # den 3, offsets -0.5, ~0, ~0.5
    assert len(f) == 9
    ws = sorted(w for _, w in f)
    np.testing.assert_allclose(ws[-1], 1.0, atol=1e-8)                      # the centre
    np.testing.assert_allclose(ws[0], math.exp(-8.0), rtol=1e-3)            # a corner: 2 (0.5)^2 / 0.25^2
    for mk, _ in f:
        r = Transform(mk).get_rotation("eman")
        assert abs(r["phi"]) < 1e-3 or abs(abs(r["phi"]) - 360.0) < 1e-3
        assert abs(((r["az"] - 30.0 + 180.0) % 360.0) - 180.0) <= 0.5 + 1e-4
        assert abs(r["alt"] - 40.0) <= 0.5 + 1e-4
        np.testing.assert_array_equal(mk[:, 3], m[:, 3])


def test_insert_with_fill_batches_by_offset():
    from eman2_amd.make3d import insert_with_fill

    class Rec:
        P = 96

        def __init__(self):
            self.calls = []

        def insert_batch(self, imgs, xf, w, defocus=None, ctf=None):
            self.calls.append((len(imgs), np.array(xf), np.array(w)))

    from eman2_amd.transform import euler_eman_matrix
    xf = np.stack([euler_eman_matrix(10.0 * i, 30.0, 5.0) for i in range(4)])
    r = Rec()
    seen = []
    insert_with_fill(r, np.zeros((4, 8, 8), np.float32), xf, np.full(4, 2.0, np.float32), fillangle=3.0,
                     progress=seen.append)
    # P = 96: step atan2(1, 48) = 1.19 deg, den = floor(3 / 1.19) = 2: a 2 x 2 grid, one batch per offset
    assert len(r.calls) == 4 and all(c[0] == 4 for c in r.calls)
    assert seen[-1] == 1.0 and len(seen) == len(r.calls)
    for _, xk, wk in r.calls:
        assert xk.shape == (4, 3, 4)
        np.testing.assert_allclose(wk, 2.0 * np.exp(-2 * 1.5 ** 2 / 0.75 ** 2), rtol=3e-4)  # offsets -1.5, 1.49999
