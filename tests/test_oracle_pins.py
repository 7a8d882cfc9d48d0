"""Independent pins of the oracle's normalisation (normalize_threed) and of its
slice comparison (do_compare_slice_work).  (CPU only.)

normalize_threed (libEM/reconstructor.cpp:315-388) turns the half-volume sums
that add_complex_at_fast (libEM/emdata_core.h:368-413) leaves behind into
data / weight.  The half-volume stores the x >= 0 half of a Hermitian
P x P x P grid; on the x = 0 and x = P/2 planes add_complex_at_fast adds a
value at (x, y, z) and its conjugate at (x, -y, -z) but the weight at one of
the two cells only, so the pairs are summed afterwards and the five
self-conjugate cells the loop skips are doubled.  What that has to produce is
fixed without restating any of it: accumulate every contribution c with
weight g on the full periodic grid at k and conj(c) with weight g at -k (a
real map's transform is Hermitian), divide, and read the x >= 0 half.  The
tests below do that with their own gridding loop (exact exp, periodic
indices) and compare with the oracle's finish(doift=False), after insertions
chosen to put weight on every self-conjugate cell and on the x = P/2 plane.
Dropping the five x2 cells or the x = P/2 pairing changes those cells by a
factor of 2 and fails here.

do_compare_slice_work (libEM/reconstructor.cpp:1259-1381, pixel_at
:1448-1461) is pinned by closed forms on volumes built by hand (constant
data and weight at the identity orientation: the dot products reduce to
lattice-point counts in an annulus of the half-plane, so the 12 A low-res
radius, its clamps, vweight / nval and the qual formula with a weight other
than 1 each fix a number), by a vectorised numpy statement of the loop on a
random volume under d2 (nearest voxel by Util::round, util.h:465-471; the
x = P/2+1 column read one element into the next row), and by the sub
invariant: agreement with sub on a volume holding A, B and C equals
agreement without sub on the volume holding B and C.
"""
import math

import numpy as np
import pytest

from eman2_amd.transform import Transform, euler_eman_matrix

f32 = np.float32


def cplx(sl):
    return sl[..., 0::2].astype(np.float64) + 1j * sl[..., 1::2].astype(np.float64)


def _mat(rows):
    m = np.zeros((3, 4), np.float32)
    m[:, :3] = np.asarray(rows, np.float64)
    return m.reshape(12)


# rows are the images of the slice's x and y axes and of the normal (v . M convention)
IDENTITY = _mat([[1, 0, 0], [0, 1, 0], [0, 0, 1]])
XZ_PLANE = _mat([[1, 0, 0], [0, 0, 1], [0, -1, 0]])          # slice y -> volume z
R2 = math.sqrt(0.5)
DIAGONAL = _mat([[R2, 0.5, 0.5], [R2, -0.5, -0.5], [0, R2, -R2]])   # slice (h, -h) -> volume (0, h, h)


# ------------------------------------------------------------ normalisation
def _periodic_weights(P, inserts, mode, corners):
    """Grid every contribution's weight onto the periodic P^3 grid at k and at
    -k (the weight of a real map's Hermitian transform is even) with an own
    loop (exact exp), and return the x >= 0 half in [z][y][x] order."""
    h = P // 2
    W = np.zeros((P, P, P))
    for sl, rot, w in inserts:
        M = Transform(rot).matrix
        for y in range(-h, h):
            for x in range(0, h + 1):
                if not corners and x * x + y * y > h * h + h:      # round(hypot) > P/2
                    continue
                rx, ry = f32(x) / f32(P), f32(y) / f32(P)
                c = [float(f32(f32(rx * M[0, d]) + f32(ry * M[1, d])) * f32(P)) for d in range(3)]
                if mode == "nearest_neighbor":
                    cells = [tuple(int(math.floor(q + 0.5)) for q in c)]
                else:
                    x0, y0, z0 = (int(math.floor(q)) for q in c)
                    if min(x0, y0, z0) < -h - 1 or max(x0, y0, z0) > h:
                        continue
                    cells = [(i, j, k) for k in (z0, z0 + 1) for j in (y0, y0 + 1) for i in (x0, x0 + 1)]
                for (i, j, k) in cells:
                    if abs(i) > h or abs(j) > h or abs(k) > h:
                        continue
                    di, dj, dk = i - c[0], j - c[1], k - c[2]
                    if mode == "gauss_2":
                        g = math.exp(-(di * di + dj * dj + dk * dk) * math.pi ** 2 / 4)
                    elif mode == "trilinear":
                        g = (1 - abs(di)) * (1 - abs(dj)) * (1 - abs(dk))
                    else:
                        g = 1.0
                    W[k % P, j % P, i % P] += g * w
                    W[-k % P, -j % P, -i % P] += g * w
    return W[:, :, :h + 1]


def _inserts(P, seed, n_random=3):
    rng = np.random.default_rng(seed)
    out = []
    from oracle import oracle as O
    for rot in [IDENTITY, XZ_PLANE, DIAGONAL] + [euler_eman_matrix(*(rng.random(3) * (360, 180, 360))).reshape(12)
                                                 for _ in range(n_random)]:
        sl = O.prep_particle(rng.standard_normal((P - 2, P - 2)).astype(np.float32), P)
        out.append((sl, rot, float(rng.uniform(0.5, 1.5))))
    return out


SPECIAL = [(0, 0, "h"), (0, "h", 0), (0, "h", "h"), ("h", 0, "h"), ("h", "h", 0)]   # (x, y, z)


@pytest.mark.parametrize("P,mode,corners,sqrtnorm", [
    (8, "trilinear", True, False), (16, "trilinear", True, False), (16, "trilinear", False, True),
    (8, "nearest_neighbor", True, False), (16, "gauss_2", True, False), (16, "gauss_2", False, False),
    (8, "gauss_2", True, True)])
def test_normalize_weights_are_the_periodic_weights(oracle, P, mode, corners, sqrtnorm):
    """After normalize_threed every stored cell's weight is the periodic
    grid's weight there (k and -k both counted), except (0,0,0) and
    (P/2,0,0), where add_complex_at_fast keeps the real part once (the
    imaginary part reset) and the weight once: half.  The data is the
    pre-normalisation sum over that weight (or its square root), 0 where
    the weight is 0."""
    h = P // 2
    ins = _inserts(P, seed=100 + P)
    v = oracle.Volume(P, "c1", mode)
    for sl, rot, w in ins:
        v.insert(sl, rot, w, corners=corners)
    pre = cplx(v.data)
    got = cplx(v.finish(False, sqrtnorm=sqrtnorm))
    W = _periodic_weights(P, ins, mode, corners)
    W[0, 0, 0] /= 2
    W[0, 0, h] /= 2
    # the table exp of gauss_2 is within 1.005 % of exp per term
    rt = 0.0102 if mode == "gauss_2" else 2e-6
    nf = v.norm_fixed.astype(np.float64)
    err = np.abs(nf - W)
    assert np.all(err <= rt * W + 1e-6), (np.unravel_index(np.argmax(err - rt * W), err.shape), err.max())
    d = np.sqrt(nf) if sqrtnorm else nf
    with np.errstate(divide="ignore", invalid="ignore"):
        want = np.where(nf != 0, pre / d, 0)
    np.testing.assert_allclose(got, want, rtol=2e-7, atol=1e-7 * np.abs(want).max())
    # the insertions above reach every cell that the pairing and the x2 fix-ups act on
    if corners:
        for (x, y, z) in SPECIAL:
            x, y, z = (h if q == "h" else q for q in (x, y, z))
            assert W[z, y, x] > 0, (x, y, z)
    assert np.count_nonzero(W[:, :, h]) >= P     # the x = P/2 plane holds weight


def test_normalize_x2_cells_single_contribution(oracle):
    """one slice pixel of value c landing exactly on a self-conjugate cell
    (trilinear weight 1 there): c and conj(c) meet in that cell, so its
    finished value is 2 Re(c) over the periodic weight (2 when no other pixel
    reaches the cell, which makes it Re(c)), never twice that."""
    P, h = 8, 4
    from oracle import oracle as O
    rng = np.random.default_rng(5)
    for rot, pix, cell in [(IDENTITY, (0, -h), (0, h, 0)), (IDENTITY, (h, -h), (h, h, 0)),
                           (XZ_PLANE, (0, -h), (0, 0, h)), (XZ_PLANE, (h, -h), (h, 0, h)),
                           (DIAGONAL, (h, -h), (0, h, h))]:
        sl = np.zeros((P, P + 2), np.float32)
        x, y = pix
        val = complex(*rng.uniform(0.5, 2.0, 2))
        sl[y % P, 2 * x], sl[y % P, 2 * x + 1] = val.real, val.imag
        v = O.Volume(P, "c1", "trilinear")
        v.insert(sl, rot, 1.0, corners=True)
        got = cplx(v.finish(False))
        cx, cy, cz = cell
        wp = _periodic_weights(P, [(sl, rot, 1.0)], "trilinear", True)[cz, cy, cx]
        assert wp >= 2
        assert abs(got[cz, cy, cx] - 2 * val.real / wp) < 1e-6 * abs(val), (cell, got[cz, cy, cx], val, wp)


# ------------------------------------------------------------ slice agreement
def _visited_identity(P):
    """pixels do_compare_slice_work visits at the identity orientation: the
    loop's x in [0, P/2+1], y in [-P/2, P/2) less the origin, kept when
    |x/P| <= 0.5 and |y/P| < 0.5"""
    h = P // 2
    return [(x, y) for y in range(-h, h) for x in range(0, h + 2)
            if (x, y) != (0, 0) and x <= h and abs(y) < h]


def _rmax(P, apix):
    """12 A expressed as a Fourier radius, clamped to [5, P/2 - 1]"""
    r = float(f32(float(f32(f32(P) * f32(apix))) / 12.0))
    return min(max(r, 5.0), P / 2 - 1)


def _volume_const(oracle, P, w0, value):
    v = oracle.Volume(P, "c1", "gauss_2")
    v.norm[:] = w0
    v.data[:, :, 0::2] = w0 * value.real
    v.data[:, :, 1::2] = w0 * value.imag
    return v


@pytest.mark.parametrize("P,apix,weight", [(48, 2.0, 0.25), (48, 1.5, 1.0), (32, 1.0, 4.0), (32, 30.0, 0.5),
                                           (64, 1.2, -1.0)])
def test_agreement_closed_form_band(oracle, P, apix, weight):
    """Volume = (1, 0) with weight w0 everywhere; slice = (1, 0) at pixel
    radius <= 6 and (0, 1) beyond.  Then every visited pixel has dt = (1, 0),
    dt[2] = w0 and
      absqual        = n6 / n                       (n6: visited pixels with r <= 6)
      absqual_lowres = n(3 < r <= 6, r < rmax) / n(3 < r < rmax)
      norm           = sqrt(power / power2) = 1     (|dt| = |slice| = 1)
      weight         = w0
      qual           = d rw / ((rw - 1) d + 1), d = absqual, rw = 1/weight (1 if weight <= 0)."""
    w0 = 2.5
    v = _volume_const(oracle, P, w0, 1.0 + 0j)
    sl = np.zeros((P, P + 2), np.float32)
    pts = _visited_identity(P)
    for y in range(-P // 2, P // 2):
        for x in range(0, P // 2 + 1):
            if math.hypot(x, y) <= 6:
                sl[y % P, 2 * x] = 1.0
            else:
                sl[y % P, 2 * x + 1] = 1.0
    q = v.agreement(sl, IDENTITY, weight, sub=False, apix=apix)
    r = [float(f32(math.hypot(x, y))) for x, y in pts]
    n = len(pts)
    n6 = sum(1 for rr in r if rr <= 6)
    rmax = _rmax(P, apix)
    band = [rr for rr in r if 3 < rr < rmax]
    lo = sum(1 for rr in band if rr <= 6) / len(band)
    d = n6 / n
    rw = 1.0 if weight <= 0 else 1.0 / weight
    assert q["reconstruct_absqual"] == pytest.approx(d, rel=1e-6)
    assert q["reconstruct_absqual_lowres"] == pytest.approx(lo, rel=1e-6)
    assert q["reconstruct_norm"] == pytest.approx(1.0, rel=1e-6)
    assert q["reconstruct_weight"] == pytest.approx(w0, rel=1e-6)
    assert q["reconstruct_qual"] == pytest.approx(d * rw / ((rw - 1) * d + 1), rel=1e-6)


def test_agreement_closed_form_scale(oracle):
    """slice = a x (volume value) over the visited pixels: absqual = 1 and
    norm = sqrt(power / power2) = 1 / a; a volume of zero weight is skipped
    pixel by pixel (dt[2] == 0), so zeroing the weight of the z = 0 plane's
    upper half leaves only the others in every sum."""
    P, a, w0 = 32, 0.37, 1.75
    v = _volume_const(oracle, P, w0, 0.6 - 0.8j)
    sl = np.zeros((P, P + 2), np.float32)
    sl[:, 0::2] = a * 0.6
    sl[:, 1::2] = a * -0.8
    q = v.agreement(sl, IDENTITY, 1.0, sub=False, apix=1.0)
    assert q["reconstruct_absqual"] == pytest.approx(1.0, rel=1e-6)
    assert q["reconstruct_norm"] == pytest.approx(1 / a, rel=1e-6)
    # rows y = 1 .. 5 of the z = 0 plane lose their weight: still absqual 1, weight w0
    v.norm[0, 1:6, :] = 0
    sl2 = sl.copy()
    sl2[1:6, :] = 7.0       # values there would break both if they were used
    q2 = v.agreement(sl2, IDENTITY, 1.0, sub=False, apix=1.0)
    assert q2["reconstruct_absqual"] == pytest.approx(1.0, rel=1e-6)
    assert q2["reconstruct_weight"] == pytest.approx(w0, rel=1e-6)


def _round_ref(v):
    """Util::round: half away from zero, truncating cast (util.h:465-471)"""
    return np.where(v < 0, (v - f32(0.5)).astype(np.float32).astype(np.int64),
                    (v + f32(0.5)).astype(np.float32).astype(np.int64))


def _agreement_np(P, data, norm, slice_, rot, syms, weight, apix):
    """vectorised statement of do_compare_slice_work on a fixed volume"""
    h = P // 2
    flat = slice_.astype(np.float32).ravel()
    ys, xs = np.mgrid[-h:h, 0:h + 2]
    keep = ~((xs == 0) & (ys == 0))
    xs, ys = xs[keep], ys[keep]
    rx = xs.astype(np.float32) / f32(P)
    ry = ys.astype(np.float32) / f32(P)
    rr = np.hypot(xs.astype(np.float64), ys.astype(np.float64)).astype(np.float32)
    rmax = _rmax(P, apix)
    tot = dict(dot=0.0, power=0.0, power2=0.0, dotlow=0.0, dlpow=0.0, dlpow2=0.0, vw=0.0, n=0)
    for s in syms:
        M = (Transform(rot) * Transform(s)).matrix.astype(np.float32)
        c = [(rx * M[0, d] + ry * M[1, d]).astype(np.float32) for d in range(3)]
        ok = (np.abs(c[0]) <= 0.5) & (np.abs(c[1]) < 0.5) & (np.abs(c[2]) < 0.5)
        xx, yy, zz = ((q[ok] * f32(P)).astype(np.float32) for q in c)
        x, y, r = xs[ok], ys[ok], rr[ok]
        idx = 2 * x + (P + 2) * np.where(y < 0, P + y, y)
        s0 = np.where(idx < flat.size, flat[np.minimum(idx, flat.size - 1)], 0).astype(np.float32)
        s1 = np.where(idx + 1 < flat.size, flat[np.minimum(idx + 1, flat.size - 1)], 0).astype(np.float32)
        x0, y0, z0 = _round_ref(xx), _round_ref(yy), _round_ref(zz)
        neg = x0 < 0
        ax = np.where(neg, -x0, x0)
        ay = np.where(neg, -y0, y0) % P
        az = np.where(neg, -z0, z0) % P
        vr = data[az, ay, 2 * ax]
        vi = np.where(neg, -data[az, ay, 2 * ax + 1], data[az, ay, 2 * ax + 1])
        nv = norm[az, ay, ax]
        use = nv != 0
        d0 = (vr[use] / nv[use]).astype(np.float32)
        d1 = (vi[use] / nv[use]).astype(np.float32)
        w = nv[use]
        s0, s1, r = s0[use], s1[use], r[use]
        dot = ((d0 * s0 + d1 * s1) * w).astype(np.float32).astype(np.float64)
        pw = ((d0 * d0 + d1 * d1) * w).astype(np.float32).astype(np.float64)
        pw2 = ((s0 * s0 + s1 * s1) * w).astype(np.float32).astype(np.float64)
        band = (r > 3) & (r < rmax)
        tot["dot"] += dot.sum()
        tot["power"] += pw.sum()
        tot["power2"] += pw2.sum()
        tot["dotlow"] += dot[band].sum()
        tot["dlpow"] += pw[band].sum()
        tot["dlpow2"] += pw2[band].sum()
        tot["vw"] += w.astype(np.float64).sum()
        tot["n"] += int(use.sum())
    d = tot["dot"] / math.sqrt(tot["power"] * tot["power2"])
    lo = tot["dotlow"] / math.sqrt(tot["dlpow"] * tot["dlpow2"]) if tot["dlpow"] * tot["dlpow2"] > 0 else tot["dotlow"]
    rw = 1.0 if weight <= 0 else 1.0 / weight
    return dict(reconstruct_norm=math.sqrt(tot["power"] / tot["power2"]), reconstruct_absqual=d,
                reconstruct_absqual_lowres=lo, reconstruct_qual=d * rw / ((rw - 1) * d + 1),
                reconstruct_weight=tot["vw"] / tot["n"]), tot["n"]


@pytest.mark.parametrize("P,sym,apix,weight", [(32, "d2", 1.9, 0.3), (48, "c1", 2.0, 2.0), (40, "c4", 0.7, 1.0)])
def test_agreement_matches_numpy_statement(oracle, P, sym, apix, weight):
    rng = np.random.default_rng(P)
    v = oracle.Volume(P, sym, "gauss_2")
    v.data[:] = rng.standard_normal(v.data.shape).astype(np.float32)
    v.norm[:] = rng.uniform(0.2, 3.0, v.norm.shape).astype(np.float32)
    v.norm[rng.random(v.norm.shape) < 0.1] = 0           # cells without weight are skipped
    sl = rng.standard_normal((P, P + 2)).astype(np.float32)
    rot = euler_eman_matrix(*(rng.random(3) * (360, 180, 360))).reshape(12)
    q = v.agreement(sl, rot, weight, sub=False, apix=apix)
    ref, n = _agreement_np(P, v.data, v.norm, sl, rot, oracle.sym_get(sym), weight, apix)
    assert n > 100
    for k in ref:
        assert q[k] == pytest.approx(ref[k], rel=2e-5, abs=1e-6), (k, q[k], ref[k])


@pytest.mark.parametrize("mode", ["gauss_2", "trilinear", "gauss_var"])
def test_agreement_sub_removes_own_slice(oracle, mode):
    """agreement(A, sub=True) on {A, B, C} == agreement(A, sub=False) on {B, C};
    the volume is restored afterwards"""
    P = 32
    rng = np.random.default_rng(41)
    sls = [oracle.prep_particle(rng.standard_normal((24, 24)).astype(np.float32), P) for _ in range(3)]
    rots = [euler_eman_matrix(*(rng.random(3) * (360, 180, 360))).reshape(12) for _ in range(3)]
    full = oracle.Volume(P, "c1", mode)
    rest = oracle.Volume(P, "c1", mode)
    # a seeded base under both (every cell weighted): a cell reached only by the
    # removed slice would keep the rounding residue of its +w - w sums instead
    # of 0, in the reference as well, and count in nval
    base_d = rng.standard_normal(full.data.shape).astype(np.float32)
    base_w = rng.uniform(0.5, 1.5, full.norm.shape).astype(np.float32)
    for v in (full, rest):
        v.data[:] = base_d
        v.norm[:] = base_w
    for i, (s, r) in enumerate(zip(sls, rots)):
        full.insert(s, r, 0.7)
        if i:
            rest.insert(s, r, 0.7)
    before = full.data.copy()
    a = full.agreement(sls[0], rots[0], 0.7, sub=True, apix=1.3)
    b = rest.agreement(sls[0], rots[0], 0.7, sub=False, apix=1.3)
    for k in a:
        assert a[k] == pytest.approx(b[k], rel=2e-3, abs=1e-4), (k, a[k], b[k])
    assert np.abs(full.data - before).max() <= 1e-5 * np.abs(before).max()
