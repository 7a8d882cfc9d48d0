"""CPU replay of the batched agreement's own-plane gather enumeration
(gather_own_offsets / own_offset_mask / ap_key in insert_kernels.hip): for
random orientations, the in-plane offsets flagged for a slice reach every
slice pixel whose 2x2x2 stencil (origin floor(p), the insertion's cell set,
libEM/reconstructor_tools.cpp:128-167) holds a compare pixel's nearest cell,
the mirrored cell of a compare pixel at x >= 3 is reached by no pixel of
the slice, and the packed floor keys' digit test equals the floor test.
(CPU only.)"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

f32 = np.float32
DIGITS = 0x100401


def key(x0, y0, z0):
    return (x0 + 512) | ((y0 + 512) << 10) | ((z0 + 512) << 20)


def positions(m, xs, ys, P):
    """pixel_coords: (fl(x/P) m0 + fl(y/P) m1) * P in float32, no contraction"""
    rx = (xs.astype(f32) / f32(P)).astype(f32)
    ry = (ys.astype(f32) / f32(P)).astype(f32)
    out = []
    for d in range(3):
        s = (rx * m[0, d]).astype(f32) + (ry * m[1, d]).astype(f32)
        out.append((s.astype(f32) * f32(P)).astype(f32))
    return out


def round_ref(v):
    v = v.astype(f32)
    return np.where(v < 0, np.trunc((v - f32(0.5)).astype(f32)), np.trunc((v + f32(0.5)).astype(f32))).astype(int)


def offset_mask(m):
    mask = []
    for k in range(25):
        ox, oy = f32(k % 5 - 2), f32(k // 5 - 2)
        d = [abs(f32(f32(ox * m[0, a]) + f32(oy * m[1, a]))) for a in range(3)]
        if max(d) < f32(1.51):
            mask.append((k % 5 - 2, k // 5 - 2))
    return mask


@pytest.mark.parametrize("P,seed", [(96, 1), (128, 2), (96, 3), (64, 4)])
def test_offsets_cover_every_source_and_keys_match_floors(P, seed):
    H = P // 2
    rng = np.random.default_rng(seed)
    for rot in Rotation.random(4, random_state=seed).as_matrix():
        m = rot.astype(f32)
        # slice pixels of the insertion loop (radius cut, origin range)
        X, Y = np.meshgrid(np.arange(0, H + 1), np.arange(-H, H))
        X, Y = X.ravel(), Y.ravel()
        keep = X * X + Y * Y <= H * H + H
        X, Y = X[keep], Y[keep]
        px, py, pz = positions(m, X, Y, P)
        fl = np.stack([np.floor(px), np.floor(py), np.floor(pz)], 1).astype(int)
        ok = (fl.min(1) >= -H - 1) & (fl.max(1) <= H)
        X, Y, fl = X[ok], Y[ok], fl[ok]
        by_floor = {}
        for x, y, f in zip(X, Y, map(tuple, fl)):
            by_floor.setdefault(f, []).append((x, y))
        floor_of = {(x, y): f for x, y, f in zip(X, Y, map(tuple, fl))}

        def sources(c):
            out = []
            for d in range(8):
                out += by_floor.get((c[0] - (d & 1), c[1] - ((d >> 1) & 1), c[2] - (d >> 2)), [])
            return set(out)

        offs = offset_mask(m)
        # compare pixels: x in [0, P/2 + 1], y in [-P/2, P/2), inside the cube, not the origin
        ex = rng.integers(0, H + 2, 3000)
        ey = rng.integers(-H, H, 3000)
        ex = np.concatenate([ex, np.repeat(np.arange(3), 2 * H)])
        ey = np.concatenate([ey, np.tile(np.arange(-H, H), 3)])
        cx, cy, cz = positions(m, ex, ey, P)
        for x, y, a, b, c in zip(ex, ey, cx, cy, cz):
            if (x == 0 and y == 0) or abs(a) > H or abs(b) >= H or abs(c) >= H:
                continue
            C = tuple(int(v) for v in round_ref(np.array([a, b, c])))
            cand = {(x + ox, y + oy) for ox, oy in offs}
            assert sources(C) <= cand, (C, x, y)
            mirror = sources(tuple(-v for v in C))
            if x >= 3:
                assert not mirror, (C, x, y)
            else:
                assert mirror <= {(-x + ox, -y + oy) for ox, oy in offs}, (C, x, y)
            # the packed key digit test against the floor test, for every candidate in the slice
            kc = key(*C)
            for q in cand:
                f = floor_of.get(q)
                if f is None:
                    continue
                dig = ((kc - key(*f)) & 0xFFFFFFFF) & ~DIGITS & 0xFFFFFFFF
                inside = all(0 <= C[i] - f[i] <= 1 for i in range(3))
                assert (dig == 0) == inside, (C, q, f)
