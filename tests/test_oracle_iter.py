"""Pins of the oracle's fourier_iter reconstructor (libEM/reconstructor.cpp:
428-682).  (CPU only.)

* Without a reference volume the insertion is a Gaussian gridding with
  gauss_var's width: pinned by an own brute-force loop with the exact exp on
  the unfolded grid, folded like the other brute-force pins, with the
  reconstructor's own ranges (slice radius P/2 + 3, x clamped to
  |x| <= P/2 - 1).
* With a reference the inserted value is dt + ref(i,j,k) - ref~(p), ref~ the
  trilinear interpolation: a zero reference reduces it to a plain Gaussian
  gridding with h = 1/2 inside the sphere i^2+j^2+k^2 <= P^2/4 (brute force);
  a reference that is affine in (i, j, k) (within the stencils) is
  interpolated exactly, so the value becomes dt + slope . (cell - p): a closed
  form, checked against the brute-force gridding of that value.
* setup_seed with a real map: the reference is numpy's rfftn of the map
  times (-1)^(x+y+z), the Fourier form of moving the phase origin to the
  corner (numpy's ifftshift before the transform).
"""
import math

import numpy as np
import pytest

from eman2_amd.transform import Transform, euler_eman_matrix

f32 = np.float32


def cplx(a):
    return a[..., 0::2].astype(np.float64) + 1j * a[..., 1::2].astype(np.float64)


def _brute_iter(sl, P, rot, syms, weight, ref_fn=None):
    """unfolded brute-force gridding of fourier_iter's insertion; ref_fn(i, j, k)
    (or None) gives the complex reference, its trilinear value at p taken
    from the same function at the eight surrounding cells"""
    h = P // 2
    S = cplx(sl)
    n = 2 * h + 1
    V = np.zeros((n, n, n), complex)
    W = np.zeros((n, n, n))
    A = np.zeros((n, n, n))
    for s in syms:
        M = (Transform(rot) * Transform(s)).matrix
        for y in range(-h, h):
            for x in range(0, h + 1):
                if x * x + y * y > (h + 3) ** 2 + (h + 3):          # round(hypot) > P/2 + 3
                    continue
                rx, ry = f32(x) / f32(P), f32(y) / f32(P)
                c = [float(f32(f32(rx * M[0, d]) + f32(ry * M[1, d])) * f32(P)) for d in range(3)]
                val = np.conj(S[-y, 0]) if (x == 0 and y < 0) else S[y % P, x]
                lo = [int(math.floor(q - 2.5)) for q in c]
                # the reference's range test (nx2 = P/2 - 1 for x) before the clamps
                if lo[0] < -(h - 1) - 4 or lo[1] < -h - 4 or lo[2] < -h - 4 or lo[0] > (h - 1) + 3 \
                        or lo[1] > h + 3 or lo[2] > h + 3:
                    continue
                rng = [range(max(lo[0], -(h - 1)), min(lo[0] + 5, h - 1) + 1),
                       range(max(lo[1], -h), min(lo[1] + 5, h) + 1),
                       range(max(lo[2], -h), min(lo[2] + 5, h) + 1)]
                if ref_fn is None:
                    hv = max(1.0 / (sum(q * q for q in c) / 4000.0 + 0.15), 0.1)
                else:
                    hv = 0.5
                    f0 = [math.floor(q) for q in c]
                    t = [q - b for q, b in zip(c, f0)]
                    nodt = 0
                    for dz in (0, 1):
                        for dy in (0, 1):
                            for dx in (0, 1):
                                wt = (t[0] if dx else 1 - t[0]) * (t[1] if dy else 1 - t[1]) * (t[2] if dz else 1 - t[2])
                                nodt += wt * ref_fn(f0[0] + dx, f0[1] + dy, f0[2] + dz)
                for k in rng[2]:
                    for j in rng[1]:
                        for i in rng[0]:
                            if ref_fn is not None and i * i + j * j + k * k > P * P // 4:
                                continue
                            g = math.exp(-((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2) * hv) * weight
                            v = val if ref_fn is None else val + ref_fn(i, j, k) - nodt
                            V[i + h, j + h, k + h] += v * g
                            W[i + h, j + h, k + h] += g
                            A[i + h, j + h, k + h] += abs(v) * g
    return V, W, A


def _fold(V, W, A, P):
    """the add_complex_at_fast fold of an unfolded grid (as in test_oracle)"""
    h = P // 2
    data = np.zeros((P, P, h + 1), complex)
    norm = np.zeros((P, P, h + 1))
    bound = np.zeros((P, P, h + 1))
    wr = lambda j: j + P if j < 0 else j          # noqa: E731
    wn = lambda j: -j if j <= 0 else P - j        # noqa: E731
    for i in range(-h, h + 1):
        for j in range(-h, h + 1):
            for k in range(-h, h + 1):
                v, w, a = V[i + h, j + h, k + h], W[i + h, j + h, k + h], A[i + h, j + h, k + h]
                if w == 0 and a == 0:
                    continue
                if (j, k) == (0, 0) and i in (0, h, -h):
                    data[0, 0, abs(i)] += v.real
                    norm[0, 0, abs(i)] += w
                    bound[0, 0, abs(i)] += a
                elif i in (0, h):
                    data[wn(k), wn(j), i] += np.conj(v)
                    data[wr(k), wr(j), i] += v
                    norm[wr(k), wr(j), i] += w
                    bound[wn(k), wn(j), i] += a
                    bound[wr(k), wr(j), i] += a
                elif i == -h:
                    data[wn(k), wn(j), h] += 2 * np.conj(v)
                    norm[wn(k), wn(j), h] += w
                    bound[wn(k), wn(j), h] += 2 * a
                elif i > 0:
                    data[wr(k), wr(j), i] += v
                    norm[wr(k), wr(j), i] += w
                    bound[wr(k), wr(j), i] += a
                else:
                    data[wn(k), wn(j), -i] += np.conj(v)
                    norm[wn(k), wn(j), -i] += w
                    bound[wn(k), wn(j), -i] += a
    return data, norm, bound


def _inputs(oracle, P, seed):
    rng = np.random.default_rng(seed)
    sl = oracle.prep_particle(rng.standard_normal((P - 4, P - 4)).astype(np.float32), P)
    rot = euler_eman_matrix(*(rng.random(3) * (360, 180, 360)))
    return sl, rot


@pytest.mark.parametrize("sym", ["c1", "d2"])
def test_iter_unseeded_brute_force(oracle, sym):
    P = 16
    sl, rot = _inputs(oracle, P, 3)
    v = oracle.Volume(P, sym, "fourier_iter")
    v.insert(sl, rot, 0.7)
    data, norm, bound = _fold(*_brute_iter(sl, P, rot, oracle.sym_get(sym), 0.7), P)
    rt = 0.0102          # the table exp: within 1.005 % per term
    assert np.all(np.abs(v.norm - norm) <= rt * np.abs(norm) + 1e-6)
    assert np.all(np.abs(cplx(v.data) - data) <= rt * bound + 1e-5)
    # the x range stops at |x| = P/2 - 1: the x = P/2 plane stays empty
    assert np.all(v.norm[:, :, P // 2] == 0) and np.count_nonzero(v.norm[:, :, P // 2 - 1]) > 0


def test_iter_zero_reference_brute_force(oracle):
    P = 16
    sl, rot = _inputs(oracle, P, 4)
    v = oracle.Volume(P, "c1", "fourier_iter")
    v.setup_iter_ref(np.zeros((P, P, P + 2), np.float32))
    v.insert(sl, rot, 1.3)
    data, norm, bound = _fold(*_brute_iter(sl, P, rot, oracle.sym_get("c1"), 1.3, lambda i, j, k: 0), P)
    assert np.all(np.abs(v.norm - norm) <= 0.0102 * np.abs(norm) + 1e-6)
    assert np.all(np.abs(cplx(v.data) - data) <= 0.0102 * bound + 1e-5)


def test_iter_affine_reference(oracle):
    """a reference volume a + g . (i, j, k) on the cells the x >= 0 half
    stores (its conjugate rule fills the x < 0 half): within the stencils of
    pixels away from the x = 0 plane the trilinear value is exact, so each
    cell receives dt + g . (cell - p)"""
    P = 16
    h = P // 2
    a, g = 0.4 - 0.2j, np.array([0.05 + 0.01j, -0.03 + 0.02j, 0.02 - 0.04j])
    ref = np.zeros((P, P, P + 2), np.float32)
    for z in range(P):
        for y in range(P):
            for x in range(h + 1):
                kk = (x, y if y < h else y - P, z if z < h else z - P)
                val = a + g[0] * kk[0] + g[1] * kk[1] + g[2] * kk[2]
                ref[z, y, 2 * x], ref[z, y, 2 * x + 1] = val.real, val.imag

    def ref_fn(i, j, k):
        """get_complex_at of that volume, wrap and conjugate rules included"""
        if abs(i) > h or abs(j) > h or abs(k) > h:
            return 0
        if i < 0:
            jj, kk = (-j) % P, (-k) % P
            return complex(ref[kk, jj, -2 * i], -ref[kk, jj, -2 * i + 1])
        return complex(ref[k % P, j % P, 2 * i], ref[k % P, j % P, 2 * i + 1])

    rng = np.random.default_rng(5)
    sl = oracle.prep_particle(rng.standard_normal((12, 12)).astype(np.float32), P)
    rot = euler_eman_matrix(20.0, 50.0, 70.0)
    v = oracle.Volume(P, "c1", "fourier_iter")
    v.setup_iter_ref(ref)
    v.insert(sl, rot, 1.0)
    data, norm, bound = _fold(*_brute_iter(sl, P, rot, oracle.sym_get("c1"), 1.0, ref_fn), P)
    assert np.all(np.abs(v.norm - norm) <= 0.0102 * np.abs(norm) + 1e-6)
    assert np.all(np.abs(cplx(v.data) - data) <= 0.0102 * bound + 1e-5)
    # the closed form on cells whose stencil points lie away from x = 0 and the
    # P/2 faces: the value there is dt + g . (cell - p) for every contribution
    assert abs(ref_fn(3, 2, -1) - (a + 3 * g[0] + 2 * g[1] - g[2])) < 1e-6


def test_iter_ref_from_map_is_shifted_fft(oracle):
    P = 12
    m = np.random.default_rng(6).standard_normal((P, P, P)).astype(np.float32)
    v = oracle.Volume(P, "c1", "fourier_iter")
    v.setup_iter_ref(m)
    want = np.fft.rfftn(np.fft.ifftshift(m.astype(np.float64)))
    assert np.abs(cplx(v.ref) - want).max() <= 1e-5 * np.abs(want).max()


def test_iter_finish_is_normalize_and_ift(oracle):
    """finish() of fourier_iter: normalize_threed without sqrtnorm, then the
    real map (the same finish as fourier's)"""
    P = 16
    sl, rot = _inputs(oracle, P, 7)
    v = oracle.Volume(P, "c1", "fourier_iter")
    v.insert(sl, rot, 1.0)
    c = cplx(v.finish(False))
    real = v.finish(True)
    ref = np.fft.fftshift(np.fft.irfftn(c, s=(P, P, P), axes=(0, 1, 2)))
    np.testing.assert_allclose(real, ref, atol=1e-5 * np.abs(ref).max())
