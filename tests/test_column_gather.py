"""CPU emulation of the column-gather insertion (k_column_insert and
k_merge_partials in insert_kernels.hip): the per-lane parallelogram
enumeration of a (slice, brick) pair must hand every (pixel, in-bounds
stencil cell) contribution of the reference loop
(libEM/reconstructor.cpp:1166-1209) to exactly one (brick, column, cell)
slot, and nothing else.  Restates the kernel's index arithmetic in Python
(float32 where the kernel computes in float32).  (CPU only.)"""
import math

import numpy as np
import pytest

from eman2_amd.transform import euler_eman_matrix
from test_binning import bricks_for_slice

f32 = np.float32
BS = 8
DLT = 2e-3
EPS = 1e-3


def axes(a):
    return a, (1 if a == 0 else 0), (1 if a == 2 else 2)


def column_setup(m, lo, A):
    """Per (slice, brick) set-up of k_column_insert (float32, as the kernel)."""
    _, B1, B2 = axes(A)
    u1, u2, v1, v2 = f32(m[0, B1]), f32(m[0, B2]), f32(m[1, B1]), f32(m[1, B2])
    D = f32(f32(u1 * v2) - f32(u2 * v1))
    iD = f32(f32(1.0) / D)
    C1, C2 = f32(lo[B1] + 3.5), f32(lo[B2] + 3.5)
    xc0 = int(np.rint(f32(f32(f32(v2 * C1) - f32(v1 * C2)) * iD)))
    yc0 = int(np.rint(f32(f32(f32(u1 * C2) - f32(u2 * C1)) * iD)))
    P01 = f32(f32(f32(xc0) * u1) + f32(f32(yc0) * v1))
    P02 = f32(f32(f32(xc0) * u2) + f32(f32(yc0) * v2))
    rows_y = f32(abs(u1) + abs(u2)) <= f32(abs(v1) + abs(v2))
    a1, a2, b1, b2 = (u1, u2, v1, v2) if rows_y else (v1, v2, u1, u2)
    iDp = iD if rows_y else -iD
    return dict(xc0=xc0, yc0=yc0, P01=P01, P02=P02, rows_y=rows_y, a1=a1, a2=a2, b1=b1, b2=b2, iDp=iDp)


def column_pixels(s, c1, c2, H):
    """Candidate pixels of column (c1, c2): the kernel's per-lane loop."""
    e1 = f32(f32(c1) - s["P01"])
    e2 = f32(f32(c2) - s["P02"])
    a1, a2, iDp = s["a1"], s["a2"], s["iDp"]
    rmid = f32(f32(f32(a1 * e2) - f32(a2 * e1)) * iDp)
    rr = f32(f32(f32(abs(a1) + abs(a2)) * f32(1 + DLT)) * abs(iDp))
    Ra = math.ceil(f32(f32(rmid - rr) - f32(EPS)))
    Rb = math.floor(f32(f32(rmid + rr) + f32(EPS)))
    xc0, yc0 = s["xc0"], s["yc0"]
    if s["rows_y"]:
        Ra, Rb = max(Ra, -H - yc0), min(Rb, H - 1 - yc0)
        Smin, Smax = -xc0, H - xc0
    else:
        Ra, Rb = max(Ra, -xc0), min(Rb, H - xc0)
        Smin, Smax = -H - yc0, H - 1 - yc0
    cons = []
    for e, a, b in ((e1, a1, s["b1"]), (e2, a2, s["b2"])):
        if abs(a) >= 0.25:
            ia = f32(f32(1.0) / a)
            cons.append((f32(f32(e - f32(1 + DLT)) * ia), f32(f32(e + f32(1 + DLT)) * ia), f32(b * ia)))
    assert cons, "at least one in-row constraint must be usable"
    out = []
    for R in range(Ra, Rb + 1):
        lo, hi = f32(Smin), f32(Smax)
        for k0, k1, sl in cons:
            t0 = f32(k0 - f32(f32(R) * sl))
            t1 = f32(k1 - f32(f32(R) * sl))
            lo, hi = max(lo, min(t0, t1)), min(hi, max(t0, t1))
        Sa = max(Smin, math.ceil(f32(lo - f32(EPS))))
        Sb = min(Smax, math.floor(f32(hi + f32(EPS))))
        for S in range(Sa, Sb + 1):
            out.append((xc0 + S, yc0 + R) if s["rows_y"] else (xc0 + R, yc0 + S))
    return out


def exact_coords(x, y, m, P):
    rx, ry = f32(x) / f32(P), f32(y) / f32(P)
    return [f32(f32(f32(rx * m[0, d]) + f32(ry * m[1, d])) * f32(P)) for d in range(3)]


def gathered(m, P, corners):
    """(pixel, cell) contributions the column kernel adds, as a multiset."""
    H = P // 2
    nb = (P + 1 + BS - 1) // BS
    n = np.cross(m[0, :3].astype(np.float64), m[1, :3].astype(np.float64))
    A = int(np.argmax(np.abs(n)))
    _, B1, B2 = axes(A)
    out = {}
    visits = []
    for bc in bricks_for_slice(m, H, BS, nb, corners):
        lo = [-H + BS * b for b in bc]
        s = column_setup(m, lo, A)
        per_lane = []
        for lane in range(64):
            c1, c2 = lo[B1] + (lane & 7), lo[B2] + (lane >> 3)
            if c1 > H or c2 > H:
                per_lane.append(0)
                continue
            pix = column_pixels(s, c1, c2, H)
            assert len(pix) == len(set(pix))
            per_lane.append(len(pix))
            for x, y in pix:
                if not corners and x * x + y * y > H * H + H:
                    continue
                c = exact_coords(x, y, m, P)
                f = [int(math.floor(q)) for q in c]
                if min(f) < -H - 1 or max(f) > H:
                    continue
                if not (0 <= c1 - f[B1] <= 1 and 0 <= c2 - f[B2] <= 1):
                    continue
                t = f[A] - lo[A]
                if t < -1 or t > 7:
                    continue
                for dk in (0, 1):
                    slot = t + 1 + dk
                    ca = f[A] + dk
                    if slot in (0, 9) or ca > H:
                        continue  # pad slots and cells past P/2 are never merged
                    cell = [0, 0, 0]
                    cell[A], cell[B1], cell[B2] = ca, c1, c2
                    key = ((x, y), tuple(cell))
                    out[key] = out.get(key, 0) + 1
        visits.append(max(per_lane))
    return out, visits


def reference_contributions(m, P, corners):
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


ANGLES = [(0.0, 0.0, 0.0), (90.0, 90.0, 0.0), (0.0, 90.0, 90.0), (45.0, 45.0, 45.0), (35.26, 45.0, 0.0),
          (0.0, 54.7356, 45.0)]


@pytest.mark.parametrize("P", [16, 20, 32])
def test_each_contribution_gathered_exactly_once(P):
    rng = np.random.default_rng(P)
    angles = ANGLES + [tuple(rng.random(3) * (360, 180, 360)) for _ in range(5)]
    for ang in angles:
        m = euler_eman_matrix(*ang)
        for corners in (False, True):
            got, _ = gathered(m, P, corners)
            want = reference_contributions(m, P, corners)
            assert set(got) == want, (ang, corners, len(set(got) ^ want))
            assert max(got.values()) == 1, (ang, corners)


def test_lane_loop_length_bound():
    """The per-lane candidate count stays small (the wave's loop trip count)."""
    rng = np.random.default_rng(5)
    for _ in range(6):
        m = euler_eman_matrix(*(rng.random(3) * (360, 180, 360)))
        _, visits = gathered(m, 48, False)
        assert max(visits) <= 24
