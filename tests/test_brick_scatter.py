"""CPU emulation of the brick scatter insertion (k_scatter_insert in
insert_kernels.hip): for every (slice, brick) pair the binning lists, the
kernel's float32 row / x-range enumeration must include every pixel whose
stencil origin lies in the brick's [lo, lo+7]^3, so that each (pixel,
in-bounds stencil cell) contribution of the reference loop
(libEM/reconstructor.cpp:1166-1209) is added exactly once, by the brick
owning the pixel's origin (cells [lo, lo+8]^3, the high faces shared with
the next bricks in the merge), and each pixel is enumerated at most once per
brick.  Restates the kernel's index arithmetic in Python.  (CPU only.)"""
import math

import numpy as np
import pytest

from eman2_amd.transform import euler_eman_matrix
from test_binning import brick_count, bricks_for_slice

f32 = np.float32
BS = 8
DLT = f32(2e-3)
EPS = f32(1e-3)
RAD = f32(6.95)


ANGLES = [(0.0, 0.0, 0.0), (90.0, 90.0, 0.0), (0.0, 90.0, 90.0), (45.0, 45.0, 45.0), (35.26, 45.0, 0.0),
          (0.0, 54.7356, 45.0)]


def exact_coords(x, y, m, P):
    """Pixel position as the reference computes it (FP32, reconstructor.cpp:1171-1196)."""
    rx, ry = f32(x) / f32(P), f32(y) / f32(P)
    return [f32(f32(f32(rx * m[0, d]) + f32(ry * m[1, d])) * f32(P)) for d in range(3)]


def reference_contributions(m, P, corners):
    """(pixel, unfolded cell) pairs the reference loop adds (cells with
    |coordinate| <= P/2, as add_complex_at_fast keeps them)."""
    H = P // 2
    out = set()
    for y in range(-H, H):
        for x in range(0, H + 1):
            if not corners and x * x + y * y > H * H + H:
                continue
            c = exact_coords(x, y, m, P)
            x0, y0, z0 = (int(math.floor(q)) for q in c)
            if min(x0, y0, z0) < -H - 1 or max(x0, y0, z0) > H:
                continue
            for k in (z0, z0 + 1):
                for j in (y0, y0 + 1):
                    for i in (x0, x0 + 1):
                        if abs(i) > H or abs(j) > H or abs(k) > H:
                            continue
                        out.add(((x, y), (i, j, k)))
    return out


def brick_pixels(m, lo, H, corners):
    """The pixels k_scatter_insert visits for one (slice, brick) pair."""
    mm = [f32(m[0, 0]), f32(m[0, 1]), f32(m[0, 2]), f32(m[1, 0]), f32(m[1, 1]), f32(m[1, 2])]
    cx, cy, cz = f32(lo[0] + 4.0), f32(lo[1] + 4.0), f32(lo[2] + 4.0)
    sc = f32(f32(f32(cx * mm[0]) + f32(cy * mm[1])) + f32(cz * mm[2]))
    tc = f32(f32(f32(cx * mm[3]) + f32(cy * mm[4])) + f32(cz * mm[5]))
    xc0, yc0 = int(np.rint(sc)), int(np.rint(tc))
    dya, dyb = math.ceil(f32(f32(tc - f32(yc0)) - RAD)), math.floor(f32(f32(tc - f32(yc0)) + RAD))
    dxa, dxb = math.ceil(f32(f32(sc - f32(xc0)) - RAD)), math.floor(f32(f32(sc - f32(xc0)) + RAD))
    assert dyb - dya + 1 <= 64
    out = []
    for lane in range(64):
        dy = dya + lane
        y = yc0 + dy
        if dy > dyb or y < -H or y > H - 1:
            continue
        dyf = f32(dy)
        xl, xh = f32(dxa), f32(dxb)
        for d in range(3):
            ud, vd = mm[d], mm[3 + d]
            if abs(ud) >= 0.05:
                P0 = f32(f32(f32(xc0) * ud) + f32(f32(yc0) * vd))
                iu = f32(f32(1.0) / ud)
                t0 = f32(f32(f32(f32(f32(lo[d]) - P0) - DLT) - f32(dyf * vd)) * iu)
                t1 = f32(f32(f32(f32(f32(lo[d] + BS) - P0) + DLT) - f32(dyf * vd)) * iu)
                xl, xh = max(xl, min(t0, t1)), min(xh, max(t0, t1))
        a = max(math.ceil(f32(xl - EPS)), -xc0)
        b = min(math.floor(f32(xh + EPS)), H - xc0)
        if not corners:
            b = min(b, math.floor(math.sqrt(H * H + H - y * y)) - xc0)
        out.extend((xc0 + dx, y) for dx in range(a, b + 1))
    return out


def scattered(m, P, corners):
    H = P // 2
    nb = brick_count(P, BS)
    got = {}
    for bc in bricks_for_slice(m, H, BS, nb, corners):
        lo = [-H - 1 + BS * b for b in bc]
        pix = brick_pixels(m, lo, H, corners)
        assert len(pix) == len(set(pix))
        for x, y in pix:
            c = exact_coords(x, y, m, P)
            f = [int(math.floor(q)) for q in c]
            if min(f) < -H - 1 or max(f) > H:
                continue
            if not all(lo[d] <= f[d] <= lo[d] + BS - 1 for d in range(3)):
                continue  # origin owned by another brick
            for dk in (0, 1):
                for dj in (0, 1):
                    for di in (0, 1):
                        cell = (f[0] + di, f[1] + dj, f[2] + dk)
                        if max(abs(q) for q in cell) > H:
                            continue  # cells past P/2 are never merged
                        key = ((x, y), cell)
                        got[key] = got.get(key, 0) + 1
    return got


@pytest.mark.parametrize("P", [16, 20, 32])
def test_each_contribution_scattered_exactly_once(P):
    rng = np.random.default_rng(P + 1)
    angles = ANGLES + [tuple(rng.random(3) * (360, 180, 360)) for _ in range(5)]
    for ang in angles:
        m = euler_eman_matrix(*ang)
        for corners in (False, True):
            got = scattered(m, P, corners)
            want = reference_contributions(m, P, corners)
            assert set(got) == want, (ang, corners, len(set(got) ^ want))
            assert max(got.values()) == 1, (ang, corners)
