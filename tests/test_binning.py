"""CPU emulation of the brick binning (k_bin) and of the per-(brick, slice)
pixel enumeration of k_brick_insert: every (pixel, in-bounds stencil cell)
contribution of the reference loop (libEM/reconstructor.cpp:1166-1209) must
be covered by a brick listed for the slice and by a pixel enumerated for that
brick.  Restates the kernels' index arithmetic in Python.  (CPU only.)"""
import math

import numpy as np
import pytest

from eman2_amd.transform import euler_eman_matrix


def _frame(m):
    u = m[0, :3].astype(np.float64)
    v = m[1, :3].astype(np.float64)
    n = np.cross(u, v)
    a = int(np.argmax(np.abs(n)))
    b1 = 1 if a == 0 else 0
    b2 = 1 if a == 2 else 2
    return u, v, n, a, b1, b2


def bricks_for_slice(m, H, BS, nb, corners=False):
    """k_bin: bricks a slice's pixels can reach (insert_kernels.hip)."""
    u, v, n, a, b1, b2 = _frame(m)
    out = set()
    dlt = 1e-3
    for c2 in range(nb):
        for c1 in range(nb):
            L1, U1 = -H + BS * c1 - 1 - dlt, -H + BS * c1 + BS + dlt
            L2, U2 = -H + BS * c2 - 1 - dlt, -H + BS * c2 + BS + dlt
            pas = [-(n[b1] * p1 + n[b2] * p2) / n[a] for p1 in (L1, U1) for p2 in (L2, U2)]
            lo = max(math.floor((min(pas) - dlt + H - BS) / BS) + 1, 0)
            hi = min(math.floor((max(pas) + dlt + H + 1) / BS), nb - 1)
            for ca in range(lo, hi + 1):
                bc = [0, 0, 0]
                bc[a], bc[b1], bc[b2] = ca, c1, c2
                c = np.array([-H + BS * bc[d] - 1 + 0.5 * (BS + 1) for d in range(3)])
                s, t = c @ u, c @ v
                rh = 0.8660254037844386 * (BS + 1) + 0.01
                if s < -rh:
                    continue
                if corners:
                    if not (s <= H + rh and -H - rh <= t <= H - 1 + rh):
                        continue
                elif s * s + t * t > (math.sqrt(H * H + H) + rh) ** 2:
                    continue
                out.add(tuple(bc))
    return out


def enum_pixels(m, H, BS, bc, corners=False):
    """k_brick_insert: candidate pixels of one (brick, slice) pair."""
    lo = [-H + BS * b for b in bc]
    uu = m[0, :3].astype(np.float64)
    vv = m[1, :3].astype(np.float64)
    L = [lo[d] - 1 - 1e-4 for d in range(3)]
    U = [lo[d] + BS + 1e-4 for d in range(3)]
    ymin = sum(L[d] * vv[d] if vv[d] >= 0 else U[d] * vv[d] for d in range(3))
    ymax = sum(U[d] * vv[d] if vv[d] >= 0 else L[d] * vv[d] for d in range(3))
    xmin = sum(L[d] * uu[d] if uu[d] >= 0 else U[d] * uu[d] for d in range(3))
    xmax = sum(U[d] * uu[d] if uu[d] >= 0 else L[d] * uu[d] for d in range(3))
    ya, yb = max(-H, math.ceil(ymin - 1e-3)), min(H - 1, math.floor(ymax + 1e-3))
    assert yb - ya + 1 <= 64
    pix = []
    for y in range(ya, yb + 1):
        xl, xh = max(0.0, xmin - 1e-3), min(float(H), xmax + 1e-3)
        for d in range(3):
            yv = y * vv[d]
            if abs(uu[d]) > 1e-9:
                iu = 1.0 / uu[d]
                t1, t2 = (L[d] - yv) * iu, (U[d] - yv) * iu
                xl, xh = max(xl, min(t1, t2) - 1e-3), min(xh, max(t1, t2) + 1e-3)
            elif yv < L[d] or yv > U[d]:
                xh = -1.0
        if not corners:
            xh = min(xh, math.floor(math.sqrt(H * H + H - y * y)))
        pix.extend((x, y) for x in range(math.ceil(xl), math.floor(xh) + 1))
    return pix


def coverage(P, BS, m, corners):
    H = P // 2
    nb = (P + 1 + BS - 1) // BS
    covered = {bc: set(enum_pixels(m, H, BS, bc, corners)) for bc in bricks_for_slice(m, H, BS, nb, corners)}
    f = np.float32
    missing = total = 0
    for y in range(-H, H):
        for x in range(0, H + 1):
            if not corners and x * x + y * y > H * H + H:
                continue
            rx, ry = f(x) / f(P), f(y) / f(P)
            c = [f(f(f(rx * m[0, d]) + f(ry * m[1, d])) * f(P)) for d in range(3)]
            x0, y0, z0 = (int(math.floor(q)) for q in c)
            if min(x0, y0, z0) < -H - 1 or max(x0, y0, z0) > H:
                continue
            for k in (z0, z0 + 1):
                for j in (y0, y0 + 1):
                    for i in (x0, x0 + 1):
                        if abs(i) > H or abs(j) > H or abs(k) > H:
                            continue
                        total += 1
                        bc = ((i + H) // BS, (j + H) // BS, (k + H) // BS)
                        if (x, y) not in covered.get(bc, ()):
                            missing += 1
    return missing, total


@pytest.mark.parametrize("P,BS", [(16, 8), (16, 16), (32, 16), (48, 8)])
def test_every_contribution_is_covered(P, BS):
    rng = np.random.default_rng(P * 100 + BS)
    angles = [(0.0, 0.0, 0.0), (90.0, 90.0, 0.0), (0.0, 90.0, 90.0), (45.0, 45.0, 45.0)]
    angles += [tuple(rng.random(3) * (360, 180, 360)) for _ in range(4)]
    for ang in angles:
        m = euler_eman_matrix(*ang)
        for corners in (False, True):
            miss, tot = coverage(P, BS, m, corners)
            assert tot > 0 and miss == 0, (ang, corners, miss, tot)


def test_bricks_per_column_bound():
    """At most 5 bricks per brick column along the dominant normal axis, the
    bound that sizes the brick list (engine.cpp ensure_bins)."""
    rng = np.random.default_rng(7)
    P, BS = 48, 8
    H, nb = P // 2, (P + 1 + BS - 1) // BS
    for _ in range(10):
        m = euler_eman_matrix(*(rng.random(3) * (360, 180, 360)))
        _, _, _, a, b1, b2 = _frame(m)
        cols = {}
        for bc in bricks_for_slice(m, H, BS, nb, corners=True):
            cols[(bc[b1], bc[b2])] = cols.get((bc[b1], bc[b2]), 0) + 1
        assert max(cols.values()) <= 5
