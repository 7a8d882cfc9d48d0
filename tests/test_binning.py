"""CPU emulation of the brick binning (k_bin_cols in insert_kernels.hip): brick b
owns the stencil origins [-P/2-1 + 8b, -P/2-1 + 8b + 8) per axis, and a slice
is listed under every brick that may own one of its pixels' origins
(libEM/reconstructor.cpp:1166-1209 for the pixel loop).  Restates the
kernel's index arithmetic in Python; the coverage itself is checked through
the scatter emulation in test_brick_scatter.py.  (CPU only.)"""
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
    """k_bin_cols: bricks a slice is listed under (insert_kernels.hip)."""
    u, v, n, a, b1, b2 = _frame(m)
    out = set()
    dlt = 1e-3
    for c2 in range(nb):
        for c1 in range(nb):
            L1, U1 = -H - 1 + BS * c1 - dlt, -H - 1 + BS * c1 + BS + dlt
            L2, U2 = -H - 1 + BS * c2 - dlt, -H - 1 + BS * c2 + BS + dlt
            r1, r2 = -n[b1] / n[a], -n[b2] / n[a]
            pas = [r1 * p1 + r2 * p2 for p1 in (L1, U1) for p2 in (L2, U2)]
            lo = max(math.floor((min(pas) - dlt + H + 1) / BS), 0)
            hi = min(math.floor((max(pas) + dlt + H + 1) / BS), nb - 1)
            for ca in range(lo, hi + 1):
                bc = [0, 0, 0]
                bc[a], bc[b1], bc[b2] = ca, c1, c2
                c = np.array([-H - 1 + BS * bc[d] + 0.5 * BS for d in range(3)])
                s, t = c @ u, c @ v
                rh = 0.8660254037844386 * BS + 0.01
                if s < -rh:
                    continue
                if corners:
                    if not (s <= H + rh and -H - rh <= t <= H - 1 + rh):
                        continue
                elif s * s + t * t > (math.sqrt(H * H + H) + rh) ** 2:
                    continue
                out.add(tuple(bc))
    return out


def brick_count(P, BS=8):
    """Bricks per axis over the stencil origins [-P/2-1, P/2] (engine.cpp brick_geom)."""
    return (P + 2 + BS - 1) // BS


def test_bricks_per_column_bound():
    """At most 4 bricks per brick column along the dominant normal axis
    (the plane spans < 2 BS + margin there), inside the bound of 5 that sizes
    the brick list (engine.cpp ensure_bins)."""
    rng = np.random.default_rng(7)
    P, BS = 48, 8
    H, nb = P // 2, brick_count(P, BS)
    for _ in range(10):
        m = euler_eman_matrix(*(rng.random(3) * (360, 180, 360)))
        _, _, _, a, b1, b2 = _frame(m)
        cols = {}
        for bc in bricks_for_slice(m, H, BS, nb, corners=True):
            cols[(bc[b1], bc[b2])] = cols.get((bc[b1], bc[b2]), 0) + 1
        assert max(cols.values()) <= 4
