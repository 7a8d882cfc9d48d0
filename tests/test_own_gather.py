"""CPU emulation of the batched agreement's own-contribution gather
(own_contribution / gather_cell in insert_kernels.hip): at every storage
voxel it must equal what inserting that particle alone into an empty volume
puts there (the oracle's add_complex_at_fast folding,
libEM/emdata_core.h:368-413).  (CPU only.)"""
import math

import numpy as np
import pytest

from eman2_amd.transform import Transform, euler_eman_matrix

f32 = np.float32
HALF_PI2 = f32(1.0) / f32(4.0 / (math.pi * math.pi))


def gather_cell(S, m, ci, cj, ck, P, weight, table, mode):
    H = P // 2
    n = np.cross(m[0, :3].astype(np.float64), m[1, :3].astype(np.float64))
    if abs(n @ (ci, cj, ck)) > 1.7330:
        return 0.0, 0.0, 0.0
    s = m[0, :3].astype(np.float64) @ (ci, cj, ck)
    t = m[1, :3].astype(np.float64) @ (ci, cj, ck)
    re = im = w = 0.0
    for y in range(max(-H, math.ceil(t - 1.733)), min(H - 1, math.floor(t + 1.733)) + 1):
        for x in range(max(0, math.ceil(s - 1.733)), min(H, math.floor(s + 1.733)) + 1):
            if x * x + y * y > H * H + H:
                continue
            rx, ry = f32(x) / f32(P), f32(y) / f32(P)
            c = [f32(f32(f32(rx * m[0, d]) + f32(ry * m[1, d])) * f32(P)) for d in range(3)]
            x0, y0, z0 = (int(math.floor(q)) for q in c)
            if min(x0, y0, z0) < -H - 1 or max(x0, y0, z0) > H:
                continue
            if not (0 <= ci - x0 <= 1 and 0 <= cj - y0 <= 1 and 0 <= ck - z0 <= 1):
                continue
            dx, dy, dz = f32(ci) - c[0], f32(cj) - c[1], f32(ck) - c[2]
            if mode == "gauss_2":
                r = f32(f32(dx * dx) + f32(dy * dy)) + f32(dz * dz)
                g = table[int(float(f32(r * HALF_PI2)) * 50.0 + 0.5)] * f32(weight)
            else:
                g = f32((1.0 - abs(float(dx))) * (1.0 - abs(float(dy))) * (1.0 - abs(float(dz))) * weight)
            val = np.conj(S[-y, 0]) if (x == 0 and y < 0) else S[y % P, x]
            re += val.real * g
            im += val.imag * g
            w += g
    return re, im, w


def own(S, mats, xs, ys, zs, P, weight, table, mode):
    H = P // 2
    jd = ([ys] if ys <= H else []) + ([ys - P] if ys >= H else [])
    jn = ([-ys] if ys <= H else []) + ([P - ys] if ys >= H and P - ys <= H and P - ys != -ys else [])
    kd = ([zs] if zs <= H else []) + ([zs - P] if zs >= H else [])
    kn = ([-zs] if zs <= H else []) + ([P - zs] if zs >= H and P - zs <= H and P - zs != -zs else [])

    def all_syms(i, j, k):
        tot = np.zeros(3)
        for m in mats:
            tot += gather_cell(S, m, i, j, k, P, weight, table, mode)
        return tot

    if ys == 0 and zs == 0 and xs in (0, H):
        r = all_syms(xs, 0, 0)
        if xs == H:
            r = r + all_syms(-H, 0, 0)
        return r[0], 0.0, r[2]
    ore = oim = ow = 0.0
    for j in jd:
        for k in kd:
            r, i, w = all_syms(xs, j, k)
            ore, oim, ow = ore + r, oim + i, ow + w
    for j in jn:
        for k in kn:
            if xs in (0, H):
                r, i, w = all_syms(xs, j, k)
                ore, oim = ore + r, oim - i
                if xs == H:
                    r2, i2, w2 = all_syms(-H, j, k)
                    ore, oim, ow = ore + 2 * r2, oim - 2 * i2, ow + w2
            else:
                r, i, w = all_syms(-xs, j, k)
                ore, oim, ow = ore + r, oim - i, ow + w
    return ore, oim, ow


@pytest.mark.parametrize("P,sym,mode,ang", [(16, "c1", "gauss_2", (30, 50, 70)), (16, "d2", "gauss_2", (10, 80, 200)),
                                            (16, "c1", "trilinear", (0, 0, 0)), (20, "c4", "gauss_2", (90, 90, 0))])
def test_own_gather_equals_single_insertion(oracle, P, sym, mode, ang):
    rng = np.random.default_rng(P)
    slice_ = oracle.prep_particle(rng.standard_normal((P - 4, P - 4)).astype(np.float32), P)
    rot = euler_eman_matrix(*ang)
    v = oracle.Volume(P, sym, mode)
    v.insert(slice_, rot, 0.7)
    S = slice_[:, 0::2].astype(np.float64) + 1j * slice_[:, 1::2]
    mats = [(Transform(rot) * Transform(s)).matrix for s in oracle.sym_get(sym)]
    table = oracle.fast_exp_table()
    H = P // 2
    data = v.data[..., 0::2] + 1j * v.data[..., 1::2]
    for zs in range(P):
        for ys in range(P):
            for xs in range(H + 1):
                re, im, w = own(S, mats, xs, ys, zs, P, 0.7, table, mode)
                want = data[zs, ys, xs]
                assert abs(w - v.norm[zs, ys, xs]) <= 1e-5 * max(1, abs(v.norm[zs, ys, xs])), (xs, ys, zs)
                assert abs(re - want.real) <= 1e-4 and abs(im - want.imag) <= 1e-4, (xs, ys, zs, re, im, want)
