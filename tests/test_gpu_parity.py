"""GPU parity: the HIP path (through the C ABI) against the CPU oracle and the
committed golden fixtures.  Bars (tolerances written here):
  prep / slices        max|a-b| <= 2e-5 * max|b|
  weights              per-voxel relative error <= 1e-4
  complex sums, maps   max|a-b| <= 1e-4 * max|b|   and FSC >= 0.9999 to Nyquist
(FP32 atomics make the summation order differ from the oracle's.)"""
import os
import threading

import numpy as np
import pytest

from eman2_amd import synthetic as S
from eman2_amd.metrics import fsc, max_rel_err, weight_rel_err
from eman2_amd.transform import Transform, euler_eman_matrix

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
CTF = dict(voltage=300.0, cs=2.7, ampcont=10.0, apix=1.0)


def recon(P, sym="c1", mode="gauss_2", **kw):
    from eman2_amd import Reconstructors
    print("recon P=%d sym=%s mode=%s %s" % (P, sym, mode, kw), flush=True)
    r = Reconstructors.get("fourier_hip", dict(size=(P, P, P), sym=sym, mode=mode, **kw))
    r.setup()
    return r


def oracle_volume(O, P, sym, mode, slices, mats, weights, fp64=False):
    v = O.Volume(P, sym, mode, fp64=fp64)
    for s, m, w in zip(slices, mats, weights):
        v.insert(s, m, w)
    return v


def assert_state_close(gd, gn, od, on, tol=1e-4):
    assert weight_rel_err(gn, on) <= tol, weight_rel_err(gn, on)
    assert max_rel_err(gd, od) <= tol, max_rel_err(gd, od)


def assert_map_close(a, b, tol=1e-4, fmin=0.9999):
    assert max_rel_err(a, b) <= tol, max_rel_err(a, b)
    f = fsc(a, b)
    assert f.min() >= fmin, f.min()


# ------------------------------------------------------------------ prep
@pytest.mark.parametrize("P", [16, 96, 384, 56, 88, 104, 168, 448])
def test_preprocess_slice(oracle, P):
    img = np.random.default_rng(P).standard_normal((P, P)).astype(np.float32)
    r = recon(P)
    g = r.preprocess_slice(img).numpy()
    o = oracle.preprocess_slice(img)
    assert max_rel_err(g, o) <= 2e-5
    # with a Transform whose 2-D part is a translation, mirror and scale
    t = euler_eman_matrix(30.0, 40.0, 50.0).astype(np.float32)
    t[0, :] *= -1
    t[:, :3] *= np.float32(1.1)
    t[:2, 3] = (2.5, -1.25)
    g = r.preprocess_slice(img, t).numpy()
    tr, moved = oracle.xform2d(img, t)
    assert moved
    assert max_rel_err(g, oracle.preprocess_slice(tr)) <= 2e-5


def test_prep_ctf_sign_many_particles(oracle):
    """CTF_SIGN on every pixel of 200 particles at 256 -> 384: the sign
    follows hypot_fast, a correctly rounded sqrt (a 1-ulp sqrt flipped the
    sign of whole rings of pixels near zeros of the CTF)."""
    B, P, n = 256, 384, 200
    ph = S.phantom(200, 96.0, seed=7)
    ctf = dict(voltage=300.0, cs=2.7, ampcont=10.0, bfactor=80.0)
    imgs, _, dfs = S.particles(ph, n, B, seed=8, apix=1.0, snr=0.05, ctf=ctf)
    g = recon(P).prep_batch(imgs, defocus=dfs, ctf=dict(ctf, apix=1.0))
    for i in range(n):
        o = oracle.prep_particle(imgs[i], P, dict(ctf, apix=1.0, defocus=float(dfs[i])))
        assert max_rel_err(g[i], o) <= 2e-5, i


@pytest.mark.parametrize("B,P,ctf", [(64, 96, False), (64, 96, True), (63, 96, True), (96, 96, False),
                                     (256, 384, True), (192, 288, True), (400, 600, False), (320, 480, True),
                                     (255, 384, True), (97, 192, False), (100, 100, True), (2, 96, True),
                                     (3, 96, False), (288, 288, True)])
def test_prep_batch(oracle, B, P, ctf):
    rng = np.random.default_rng(B + P)
    n = 5
    imgs = (rng.standard_normal((n, B, B)) + 3.0).astype(np.float32)
    dfs = rng.uniform(0.8, 2.5, n).astype(np.float32) if ctf else None
    r = recon(P)
    g = r.prep_batch(imgs, defocus=dfs, ctf=CTF if ctf else None)
    for i in range(n):
        c = dict(CTF, defocus=float(dfs[i])) if ctf else None
        o = oracle.prep_particle(imgs[i], P, c)
        assert max_rel_err(g[i], o) <= 2e-5, (i, max_rel_err(g[i], o))


def test_prep_batch_in_sub_batches():
    """A batch larger than the prep intermediate (2,048 particles at the pad's
    own box) runs in sub-batches: the same slices as smaller calls."""
    B = P = 96
    n = 5000
    rng = np.random.default_rng(5)
    imgs = (rng.standard_normal((n, B, B)) + 1.0).astype(np.float32)
    dfs = rng.uniform(0.8, 2.5, n).astype(np.float32)
    r = recon(P)
    g = r.prep_batch(imgs, defocus=dfs, ctf=CTF)
    for a, b in ((0, 2500), (2500, n)):
        h = r.prep_batch(imgs[a:b], defocus=dfs[a:b], ctf=CTF)
        np.testing.assert_array_equal(g[a:b], h)


# ------------------------------------------------------------------ golden fixture
@pytest.mark.parametrize("sym", ["c1", "d2"])
@pytest.mark.parametrize("mode", ["gauss_2", "trilinear"])
def test_golden_insert_slice(sym, mode):
    from eman2_amd.reconstructor import EMData
    d = np.load(os.path.join(GOLD, "recon_p16.npz"))
    P = 16
    r = recon(P, sym, mode)
    for s, m, w in zip(d["slices"], d["mats"], d["weights"]):
        r.insert_slice(EMData(s, True, {"reconstruct_preproc": 1}), Transform(m), float(w))
    gd, gn = r.get_state()
    key = "%s_%s" % (sym, mode)
    assert_state_close(gd, gn, d[key + "_data"], d[key + "_norm"], tol=1e-5)
    assert_map_close(r.finish(True, return_numpy=True), d[key + "_map"], tol=1e-5)


@pytest.mark.parametrize("sym", ["c1", "d2"])
@pytest.mark.parametrize("mode", ["gauss_2", "trilinear"])
@pytest.mark.parametrize("wpb", [1, 4])
def test_golden_batch(sym, mode, wpb):
    d = np.load(os.path.join(GOLD, "recon_p16.npz"))
    P = 16
    r = recon(P, sym, mode, waves_per_block=wpb, slices_per_item=2)
    r.insert_batch(d["imgs"], list(d["mats"]), d["weights"])
    gd, gn = r.get_state()
    key = "%s_%s" % (sym, mode)
    assert_state_close(gd, gn, d[key + "_data"], d[key + "_norm"], tol=1e-5)
    assert_map_close(r.finish(True, return_numpy=True), d[key + "_map"], tol=1e-5)


# ------------------------------------------------------------------ config-1 shaped parity
def config1_inputs(n, ctf=False, B=64, seed=1):
    ph = S.phantom(50, 24.0, seed=seed)
    imgs, mats, dfs = S.particles(ph, n, B, seed=seed + 1, apix=1.5, snr=0.1,
                                  ctf=dict(CTF) if ctf else None)
    return imgs, mats, dfs


@pytest.mark.parametrize("wpb,spi", [(4, 64), (1, 16), (2, 5)])
def test_batch_config1(oracle, wpb, spi):
    P = 96
    imgs, mats, _ = config1_inputs(300)
    w = np.ones(len(imgs), np.float32)
    r = recon(P, waves_per_block=wpb, slices_per_item=spi, chunk=128)
    r.insert_batch(imgs, list(mats), w)
    gd, gn = r.get_state()
    slices = [oracle.prep_particle(im, P) for im in imgs]
    v = oracle_volume(oracle, P, "c1", "gauss_2", slices, mats, w)
    assert_state_close(gd, gn, v.data, v.norm)
    assert_map_close(r.finish(True, return_numpy=True), v.finish(True))


def test_batch_ctf(oracle):
    P = 96
    imgs, mats, dfs = config1_inputs(120, ctf=True)
    w = np.ones(len(imgs), np.float32)
    r = recon(P)
    r.insert_batch(imgs, list(mats), w, defocus=dfs, ctf=dict(CTF, apix=1.5))
    gd, gn = r.get_state()
    slices = [oracle.prep_particle(im, P, dict(CTF, apix=1.5, defocus=float(df))) for im, df in zip(imgs, dfs)]
    v = oracle_volume(oracle, P, "c1", "gauss_2", slices, mats, w)
    assert_state_close(gd, gn, v.data, v.norm)
    assert_map_close(r.finish(True, return_numpy=True), v.finish(True))


@pytest.mark.parametrize("sym,mode", [("c4", "gauss_2"), ("d2", "trilinear"), ("d7", "gauss_2"), ("d7", "trilinear"),
                                      ("c1", "trilinear"), ("tet", "gauss_2"), ("oct", "trilinear"),
                                      ("icos", "gauss_2"), ("h6:1:25.5:3", "trilinear"),
                                      ("c1", "gauss_var"), ("d2", "gauss_var"), ("c1", "gauss_5"),
                                      ("d2", "gauss_3"), ("c4", "nearest_neighbor"), ("c1", "kaiser_bessel"),
                                      ("c1", "gridding_5"), ("d2", "gridding_7"), ("icos", "gridding_5"),
                                      ("c1", "kaiser_bessel_derived")])
def test_batch_symmetry_modes(oracle, sym, mode):
    P, B = 48, 32
    ph = S.phantom(10, 10.0, seed=5)
    imgs, mats, _ = S.particles(ph, 40, B, seed=6, snr=1.0)
    w = np.linspace(0.5, 1.5, 40).astype(np.float32)
    r = recon(P, sym, mode, slices_per_item=8)
    r.insert_batch(imgs, list(mats), w)
    gd, gn = r.get_state()
    # gridding adds the full weight to all 125 cells of its stencil: under icos a voxel
    # sums ~10^4 terms of ~1, where the oracle's sequential float sum is itself 1e-4
    # off, so that case is summed in double
    fp64 = sym == "icos" and mode.startswith("gridding")
    v = oracle_volume(oracle, P, sym, mode, [oracle.prep_particle(im, P) for im in imgs], mats, w, fp64=fp64)
    assert_state_close(gd, gn, v.data, v.norm)
    assert_map_close(r.finish(True, return_numpy=True), v.finish(True))


@pytest.mark.parametrize("sym,mode", [("d2", "gauss_3"), ("c4", "kaiser_bessel"), ("c3", "nearest_neighbor"),
                                      ("c1", "gridding_5"), ("d2", "gridding_7"), ("d2", "kaiser_bessel_derived")])
def test_wide_mode_single_slice_symmetry(oracle, sym, mode):
    """insert_slice of a wide-stencil mode under a symmetry (every symmetry
    copy of the slice in one launch), then determine_slice_agreement with the
    particle's own slice removed and re-added (reconstructor.cpp:1090-1210)"""
    from eman2_amd.reconstructor import EMData
    P, B, n = 48, 32, 24
    ph = S.phantom(10, 10.0, seed=11)
    imgs, mats, _ = S.particles(ph, n, B, seed=12, snr=1.0)
    slices = [oracle.prep_particle(im, P) for im in imgs]
    r = recon(P, sym, mode)
    v = oracle.Volume(P, sym, mode)
    for sl, m in list(zip(slices, mats)) * 2:
        assert r.insert_slice(EMData(sl, True, {"reconstruct_preproc": 1}), m, 0.75) == 0
        v.insert(sl, m, 0.75)
    gd, gn = r.get_state()
    assert_state_close(gd, gn, v.data, v.norm)
    got = r.determine_slice_agreement(EMData(slices[1], True, {"reconstruct_preproc": 1}), mats[1], 0.75, True)
    ref = v.agreement(slices[1], mats[1], 0.75, sub=True, apix=1.0)
    for k in ("reconstruct_norm", "reconstruct_absqual"):
        assert abs(got[k] - ref[k]) <= 1e-4 * max(1.0, abs(ref[k])), (k, got[k], ref[k])
    # the removed slice is back: the state is unchanged by the agreement
    ad, an = r.get_state()
    assert_state_close(ad, an, v.data, v.norm)


@pytest.mark.parametrize("mode", ["gauss_2", "gauss_var"])
def test_agreement_in_the_reference_order(oracle, mode):
    """determine_slice_agreement(sub=True) in the reference's own order
    (reconstructor.cpp:1212-1257): its removal call passes -weight, the SSNR
    flag of do_insert_slice_work (reconstructor.cpp:1160-1184), so the slice
    is added with its class_ssnr weights, compared, and added again with
    +weight; without class_ssnr the reference throws
    NotExistingObjectException"""
    from eman2_amd._lib import FourierHipError
    from eman2_amd.reconstructor import EMData
    P, B, n = 48, 32, 12
    ph = S.phantom(10, 10.0, seed=21)
    imgs, mats, _ = S.particles(ph, n, B, seed=22, snr=1.0)
    slices = [oracle.prep_particle(im, P) for im in imgs]
    r = recon(P, "c1", mode)
    v = oracle.Volume(P, "c1", mode)
    for sl, m in zip(slices, mats):
        assert r.insert_slice(EMData(sl, True, {"reconstruct_preproc": 1}), m, 1.0) == 0
        v.insert(sl, m, 1.0)
    q = np.random.default_rng(23).uniform(-0.2, 2.0, 17).astype(np.float32)
    e = EMData(slices[3], True, {"reconstruct_preproc": 1, "class_ssnr": q})
    got = r.determine_slice_agreement(e, mats[3], 0.75, True, reference_sub=True)
    v.insert_ssnr(slices[3], mats[3], q)
    ref = v.agreement(slices[3], mats[3], 0.75, sub=False, apix=1.0)
    v.insert(slices[3], mats[3], 0.75)
    for k in ("reconstruct_norm", "reconstruct_absqual", "reconstruct_absqual_lowres", "reconstruct_weight"):
        assert abs(got[k] - ref[k]) <= 1e-4 * max(1.0, abs(ref[k])), (k, got[k], ref[k])
    gd, gn = r.get_state()
    assert_state_close(gd, gn, v.data, v.norm)
    with pytest.raises(FourierHipError, match="NotExistingObjectException"):
        r.determine_slice_agreement(EMData(slices[3], True, {"reconstruct_preproc": 1}), mats[3], 0.75, True,
                                    reference_sub=True)


@pytest.mark.parametrize("sym,mode", [("d2", "gauss_2"), ("c1", "trilinear"), ("d2", "gauss_var")])
def test_usessnr_batch_and_slice(oracle, sym, mode):
    """usessnr: per-radius class_ssnr weights (reconstructor.cpp:1080-1088,
    1160-1184) in the batched scatter and the single-slice plain kernel."""
    from eman2_amd.reconstructor import EMData
    P, B, n = 48, 32, 30
    ph = S.phantom(10, 10.0, seed=7)
    imgs, mats, _ = S.particles(ph, n, B, seed=8, snr=1.0)
    ss = np.random.default_rng(9).uniform(-0.2, 2.0, (n, 17)).astype(np.float32)
    slices = [oracle.prep_particle(im, P) for im in imgs]
    v = oracle.Volume(P, sym, mode)
    for sl, m, q in zip(slices, mats, ss):
        v.insert_ssnr(sl, m, q)
    r = recon(P, sym, mode, usessnr=True, slices_per_item=8)
    r.insert_batch(imgs, list(mats), None, class_ssnr=ss)
    gd, gn = r.get_state()
    assert_state_close(gd, gn, v.data, v.norm)
    r2 = recon(P, sym, mode, usessnr=True)
    for sl, m, q in zip(slices, mats, ss):
        assert r2.insert_slice(EMData(sl, True, {"reconstruct_preproc": 1, "class_ssnr": q}), m, 5.0) == 0
        # no class_ssnr on a usessnr reconstructor: weight 0, skipped
        assert r2.insert_slice(EMData(sl, True, {"reconstruct_preproc": 1}), m, 5.0) == -1
    gd2, gn2 = r2.get_state()
    assert_state_close(gd2, gn2, v.data, v.norm)
    assert_map_close(r.finish(True, return_numpy=True), v.finish(True))
    # usessnr off: class_ssnr is ignored and the weight applies
    r3 = recon(P, sym, mode)
    r3.insert_batch(imgs, list(mats), np.full(n, 0.5, np.float32), class_ssnr=ss)
    w = oracle_volume(oracle, P, sym, mode, slices, mats, np.full(n, 0.5, np.float32))
    gd3, gn3 = r3.get_state()
    assert_state_close(gd3, gn3, w.data, w.norm)


def _with_2d(mats, seed, mirror_scale=False):
    """the particles' Transforms with 2-D translations (and some x mirrors /
    scales), the part preprocess_slice applies to the image"""
    rng = np.random.default_rng(seed)
    out = []
    for i, m in enumerate(mats):
        t = np.array(m, np.float32).reshape(3, 4).copy()
        t[:2, 3] = rng.uniform(-3.0, 3.0, 2)
        if mirror_scale and i % 3 == 1:
            t[0, :] *= -1
        if mirror_scale and i % 3 == 2:
            t[:, :3] *= np.float32(1.25)
        out.append(t)
    return np.array(out, np.float32)


@pytest.mark.parametrize("mirror_scale", [False, True])
def test_2d_transforms_in_preprocess(oracle, mirror_scale):
    """preprocess_slice's translation / mirror / scale (reconstructor.cpp:1031-1035)
    on the batched, single-slice and agreement paths."""
    from eman2_amd.reconstructor import EMData
    P, B, n = 48, 32, 24
    ph = S.phantom(10, 10.0, seed=11)
    imgs, mats, _ = S.particles(ph, n, B, seed=12, snr=1.0)
    xfs = _with_2d(mats, 13, mirror_scale)
    w = np.linspace(0.5, 1.5, n).astype(np.float32)
    slices = [oracle.prep_particle(im, P, xform=t) for im, t in zip(imgs, xfs)]
    v = oracle_volume(oracle, P, "d2", "gauss_2", slices, xfs, w)
    r = recon(P, "d2", "gauss_2", slices_per_item=8)
    r.insert_batch(imgs, list(xfs), w)
    gd, gn = r.get_state()
    assert_state_close(gd, gn, v.data, v.norm)
    # raw P x P images (zero border: edge mean 0) through insert_slice
    r2 = recon(P, "d2", "gauss_2")
    o = (P - B) // 2
    for im, t, wi in zip(imgs, xfs, w):
        pad = np.zeros((P, P), np.float32)
        pad[o:o + B, o:o + B] = im - oracle.edge_mean(im)
        r2.insert_slice(EMData(pad), t, float(wi))
    gd2, gn2 = r2.get_state()
    v2 = oracle.Volume(P, "d2", "gauss_2")
    for im, t, wi in zip(imgs, xfs, w):
        pad = np.zeros((P, P), np.float32)
        pad[o:o + B, o:o + B] = im - oracle.edge_mean(im)
        v2.insert(oracle.prep_particle(pad, P, xform=t), t, float(wi))
    assert_state_close(gd2, gn2, v2.data, v2.norm)
    assert_map_close(r.finish(True, return_numpy=True), v.finish(True))


def test_2d_transforms_in_agreement(oracle):
    """agreement_batch of translated / mirrored / scaled particles against a
    well-covered volume (sparse volumes keep summation-order residues of the
    subtracted slice, which decide whether a voxel is compared)"""
    P, B, n = 32, 24, 500
    ph = S.phantom(10, 7.0, seed=45)
    imgs, mats, _ = S.particles(ph, n, B, seed=46, snr=0.5)
    xfs = _with_2d(mats, 47, True)
    w = np.ones(n, np.float32)
    r = recon(P)
    r.insert_batch(imgs, list(xfs), w)
    got = r.agreement_batch(imgs[:12], xfs[:12], w[:12], sub=True, apix=1.0)
    v = oracle_volume(oracle, P, "c1", "gauss_2", [oracle.prep_particle(im, P, xform=t) for im, t in zip(imgs, xfs)],
                      xfs, w)
    for i in range(12):
        a = v.agreement(oracle.prep_particle(imgs[i], P, xform=xfs[i]), xfs[i], 1.0, sub=True, apix=1.0)
        for j, k in enumerate(AGREE_KEYS):
            assert abs(got[i, j] - a[k]) <= 1e-4 * max(1.0, abs(a[k])), (i, k, got[i, j], a[k])


def test_2d_transform_with_ctf(oracle):
    """a CTF flip with translations, scale and mirror at a pad of the runtime
    FFT plan (48): the padded particle flipped first, then resampled
    (e2make3dpar.py:698-716), as the oracle does"""
    P, B, n = 48, 32, 16
    ph = S.phantom(5, 10.0, seed=14)
    imgs, mats, dfs = S.particles(ph, n, B, seed=15, snr=1.0, ctf=dict(CTF))
    xfs = _with_2d(mats, 16, True)
    r = recon(P)
    r.insert_batch(imgs, list(xfs), np.ones(n, np.float32), defocus=dfs, ctf=CTF)
    gd, gn = r.get_state()
    v = oracle_volume(oracle, P, "c1", "gauss_2",
                      [oracle.prep_particle(imgs[i], P, dict(CTF, defocus=float(dfs[i])), xform=xfs[i])
                       for i in range(n)], xfs, np.ones(n, np.float32))
    assert_state_close(gd, gn, v.data, v.norm)


@pytest.mark.parametrize("P,B", [(56, 40), (88, 60), (104, 64), (168, 112)])
def test_radix_7_11_13_pads(oracle, P, B):
    """good_box_sizes pads with factors 7, 11, 13 (libpyEM/EMAN2.py:660):
    prep, batched insertion and finish() with the 3-D inverse FFT"""
    ph = S.phantom(10, B / 4.0, seed=P)
    imgs, mats, _ = S.particles(ph, 30, B, seed=P + 1, snr=1.0)
    w = np.ones(30, np.float32)
    r = recon(P, "c1", "gauss_2")
    r.insert_batch(imgs, list(mats), w)
    gd, gn = r.get_state()
    v = oracle_volume(oracle, P, "c1", "gauss_2", [oracle.prep_particle(im, P) for im in imgs], mats, w)
    assert_state_close(gd, gn, v.data, v.norm)
    assert_map_close(r.finish(True, return_numpy=True), v.finish(True))


@pytest.mark.parametrize("ret_fourier", [1, 0])
@pytest.mark.parametrize("P", [48, 56])
def test_projection(oracle, ret_fourier, P):
    """projection() (reconstructor.cpp:1383-1445) of the same volume state on
    the GPU and in the oracle, with and without translations"""
    B, n = 32, 200
    ph = S.phantom(10, 10.0, seed=21)
    imgs, mats, _ = S.particles(ph, n, B, seed=22, snr=2.0)
    v = oracle_volume(oracle, P, "c1", "gauss_2", [oracle.prep_particle(im, P) for im in imgs], mats,
                      np.ones(n, np.float32))
    r = recon(P)
    r.setup_seedandweights(v.data, v.norm)
    rng = np.random.default_rng(23)
    for k in range(6):
        t = euler_eman_matrix(*(rng.random(3) * (360, 180, 360))).astype(np.float32)
        if k % 2:
            t[:2, 3] = rng.uniform(-4, 4, 2)
        g = r.projection(t, ret_fourier).numpy()
        o = v.projection(t, ret_fourier)
        assert max_rel_err(g, o) <= (1e-6 if ret_fourier else 2e-5), (k, max_rel_err(g, o))


def test_gauss_var_single_slice_and_agreement(oracle):
    """gauss_var (reconstructor_tools.cpp:382-433) through insert_slice,
    determine_slice_agreement and the batched agreement (the reference's
    remove / compare / re-add sequence per particle for this mode)"""
    from eman2_amd.reconstructor import EMData
    P, B, n = 32, 24, 60
    ph = S.phantom(8, 7.0, seed=71)
    imgs, mats, _ = S.particles(ph, n, B, seed=72, snr=1.0)
    slices = [oracle.prep_particle(im, P) for im in imgs]
    r = recon(P, "c1", "gauss_var")
    v = oracle.Volume(P, "c1", "gauss_var")
    # every particle twice: removing one copy never leaves a voxel with only the
    # rounding residue of +w-w (whose comparison is summation-order dependent)
    for sl, m in list(zip(slices, mats)) * 2:
        r.insert_slice(EMData(sl, True, {"reconstruct_preproc": 1}), m, 1.0)
        v.insert(sl, m, 1.0)
    gd, gn = r.get_state()
    assert_state_close(gd, gn, v.data, v.norm)
    got = r.determine_slice_agreement(EMData(slices[0], True, {"reconstruct_preproc": 1}), mats[0], 1.0, True)
    ref = v.agreement(slices[0], mats[0], 1.0, sub=True, apix=1.0)
    for k in ("reconstruct_norm", "reconstruct_absqual"):
        assert abs(got[k] - ref[k]) <= 1e-4 * max(1.0, abs(ref[k])), (k, got[k], ref[k])
    got = r.agreement_batch(imgs[:4], mats[:4], np.ones(4, np.float32), sub=True, apix=1.0)
    for i in range(4):
        a = v.agreement(oracle.prep_particle(imgs[i], P), mats[i], 1.0, sub=True, apix=1.0)
        for j, k in enumerate(AGREE_KEYS[:2]):
            assert abs(got[i, j] - a[k]) <= 1e-4 * max(1.0, abs(a[k])), (i, k, got[i, j], a[k])


def test_plain_matches_batch():
    P = 96
    imgs, mats, _ = config1_inputs(60)
    w = np.ones(60, np.float32)
    a = recon(P)
    a.insert_batch(imgs, list(mats), w)
    b = recon(P)
    b.insert_plain(imgs, list(mats), w)
    ad, an = a.get_state()
    bd, bn = b.get_state()
    assert_state_close(ad, an, bd, bn, tol=1e-5)


def test_insert_remove_and_zero_weight():
    P = 96
    imgs, mats, _ = config1_inputs(30)
    r = recon(P)
    r.insert_batch(imgs, list(mats), np.ones(30, np.float32))
    r.insert_batch(imgs, list(mats), -np.ones(30, np.float32))
    r.insert_batch(imgs, list(mats), np.zeros(30, np.float32))  # skipped
    r.insert_batch(imgs[:0], [], np.zeros(0, np.float32))       # empty batch
    d, n = r.get_state()
    ref = recon(P)
    ref.insert_batch(imgs, list(mats), np.ones(30, np.float32))
    rd, rn = ref.get_state()
    # rt/pyem/test_reconstructor.py:48-88 tolerance (1e-3), relative to the peak
    assert np.abs(d).max() <= 1e-3 * np.abs(rd).max()
    assert np.abs(n).max() <= 1e-3 * np.abs(rn).max()


def test_reference_fourier_reconstructor_case(oracle):
    """rt/pyem/test_reconstructor.py:94-117: three 32x32 uniform-noise images
    inserted (raw, through preprocess_slice) at (az, alt, phi) = (2.56, 1.56,
    3.56), (3.56, 2.56, 4.56), (4.56, 3.56, 5.56), gauss_2, c1, finish(True)"""
    n = 32
    imgs = np.random.default_rng(94).random((3, n, n)).astype(np.float32)
    angles = [(2.56, 1.56, 3.56), (3.56, 2.56, 4.56), (4.56, 3.56, 5.56)]
    r = recon(n)
    v = oracle.Volume(n, "c1", "gauss_2")
    for im, (az, alt, phi) in zip(imgs, angles):
        t = Transform({"type": "eman", "az": az, "alt": alt, "phi": phi})
        assert r.insert_slice(im, t, 1.0) == 0
        v.insert(oracle.preprocess_slice(im), t.matrix, 1.0)
    gd, gn = r.get_state()
    assert_state_close(gd, gn, v.data, v.norm)
    assert_map_close(r.finish(True, return_numpy=True), v.finish(True))


@pytest.mark.parametrize("mode", ["gauss_2", "trilinear", "gauss_var", "gauss_5", "gauss_3", "nearest_neighbor",
                                  "kaiser_bessel"])
def test_insert_remove_every_mode(mode):
    """the reference's insert/remove invariant over the inserter modes
    (rt/pyem/test_reconstructor.py:48-88, tolerance 1e-3): a slice inserted
    with weight 1 and again with weight -1 leaves (numerically) nothing"""
    n = 32 if mode == "kaiser_bessel" else 16
    img = np.random.default_rng(48).random((n, n)).astype(np.float32)
    for az, alt in ((0, 0), (35, 20), (120, 75)):
        r = recon(n, "c1", mode)
        t = Transform({"type": "eman", "az": az, "alt": alt, "phi": 0.0})
        r.insert_slice(img, t, 1.0)
        d1, n1 = r.get_state()
        r.insert_slice(img, t, -1.0)
        d, w = r.get_state()
        assert np.abs(n1).max() > 0
        # 1e-3, or 1e-5 of the inserted peak where that is larger (kaiser_bessel's
        # window sums to ~2000x the slice value over its 125 cells)
        tol_d = max(1e-3, 1e-5 * float(np.abs(d1).max()))
        tol_w = max(1e-3, 1e-5 * float(np.abs(n1).max()))
        assert np.abs(d).max() < tol_d and np.abs(w).max() < tol_w, (az, alt, np.abs(d).max(), np.abs(w).max())


def test_corners_and_sqrtnorm(oracle):
    P = 32
    ph = S.phantom(6, 6.0, seed=9)
    imgs, mats, _ = S.particles(ph, 30, 24, seed=10, snr=2.0)
    w = np.ones(30, np.float32)
    r = recon(P, corners=True, sqrtnorm=True)
    r.insert_batch(imgs, list(mats), w)
    gd, gn = r.get_state()
    v = oracle.Volume(P)
    for im, m in zip(imgs, mats):
        v.insert(oracle.prep_particle(im, P), m, 1.0, corners=True)
    assert_state_close(gd, gn, v.data, v.norm)
    assert_map_close(r.finish(True, return_numpy=True), v.finish(True, sqrtnorm=True))


@pytest.mark.parametrize("P,sym,mode", [(50, "c4", "gauss_var"), (90, "c1", "gauss_var"), (50, "c1", "gridding_5"),
                                        (90, "d2", "gauss_3")])
def test_wide_modes_odd_half_size(oracle, P, sym, mode):
    """the wide-stencil brick grid when P/2 is odd (P = 50, 90): the brick
    origins start at -P/2 - 4 (-P/2 - 2 for gauss_3) and the last brick is
    partly past P/2"""
    ph = S.phantom(8, 0.2 * P, seed=P)
    imgs, mats, _ = S.particles(ph, 20, int(P * 0.66) // 2 * 2, seed=P + 1, snr=1.0)
    w = np.linspace(0.5, 1.5, 20).astype(np.float32)
    r = recon(P, sym, mode, slices_per_item=8)
    r.insert_batch(imgs, list(mats), w)
    gd, gn = r.get_state()
    v = oracle_volume(oracle, P, sym, mode, [oracle.prep_particle(im, P) for im in imgs], mats, w)
    assert_state_close(gd, gn, v.data, v.norm)
    assert_map_close(r.finish(True, return_numpy=True), v.finish(True))


@pytest.mark.parametrize("mode", ["gauss_var", "gauss_3", "gridding_7", "kaiser_bessel_derived"])
def test_corners_wide_modes_batch(oracle, mode):
    """corners (reconstructor.cpp:1172-1176: the whole half-square of slice
    pixels, no radius cut) through the wide-stencil brick scatter: its
    binning takes the half-square domain and its candidate windows the
    pixels beyond the P/2 circle"""
    P = 32
    ph = S.phantom(6, 6.0, seed=13)
    imgs, mats, _ = S.particles(ph, 24, 24, seed=14, snr=2.0)
    w = np.linspace(0.5, 1.5, 24).astype(np.float32)
    r = recon(P, "d2", mode, corners=True, slices_per_item=4)
    r.insert_batch(imgs, list(mats), w)
    gd, gn = r.get_state()
    v = oracle.Volume(P, "d2", mode)
    for im, m, wi in zip(imgs, mats, w):
        v.insert(oracle.prep_particle(im, P), m, float(wi), corners=True)
    assert_state_close(gd, gn, v.data, v.norm)


def test_finish_complex_and_normout(oracle):
    P = 32
    ph = S.phantom(6, 6.0, seed=11)
    imgs, mats, _ = S.particles(ph, 20, 24, seed=12, snr=2.0)
    w = np.ones(20, np.float32)
    r = recon(P)
    r.insert_batch(imgs, list(mats), w)
    normout = np.zeros((P, P, P // 2 + 1), np.float32)
    r.set_params({"normout": normout})
    g = r.finish(False, return_numpy=True)
    assert np.array_equal(normout, r.last_norm)
    v = oracle_volume(oracle, P, "c1", "gauss_2", [oracle.prep_particle(im, P) for im in imgs], mats, w)
    o = v.finish(False)
    assert max_rel_err(g, o) <= 1e-4
    assert weight_rel_err(r.last_norm, v.norm_fixed) <= 1e-4
    with pytest.raises(RuntimeError, match="StateError"):
        r.insert_batch(imgs, list(mats), w)  # finish() released the volume


def test_setup_seed(oracle):
    P = 16
    rng = np.random.default_rng(13)
    seed = rng.standard_normal((P, P, P + 2)).astype(np.float32)
    r = recon(P)
    r.setup_seed(seed, 0.5)
    d = np.load(os.path.join(GOLD, "recon_p16.npz"))
    r.insert_batch(d["imgs"], list(d["mats"]), d["weights"])
    gd, gn = r.get_state()
    v = oracle.Volume(P)
    v.setup_seed(seed, 0.5)
    for s, m, w in zip(d["slices"], d["mats"], d["weights"]):
        v.insert(s, m, w)
    assert_state_close(gd, gn, v.data, v.norm, tol=1e-5)


def test_seed_and_weights_round_trip():
    P = 32
    imgs, mats, _ = config1_inputs(20, B=24)
    w = np.ones(20, np.float32)
    a = recon(P)
    a.insert_batch(imgs[:10], list(mats[:10]), w[:10])
    data, norm = a.get_state()
    b = recon(P)
    b.setup_seedandweights(data, norm)
    b.insert_batch(imgs[10:], list(mats[10:]), w[10:])
    c = recon(P)
    c.insert_batch(imgs, list(mats), w)
    bd, bn = b.get_state()
    cd, cn = c.get_state()
    assert_state_close(bd, bn, cd, cn, tol=1e-5)


def test_threads_share_one_reconstructor():
    """e2make3dpar.py:355-381 calls insert_slice from several threads on one
    reconstructor; the handle serialises them."""
    P = 32
    imgs, mats, _ = config1_inputs(24, B=24)
    pad = np.zeros((24, P, P), np.float32)
    pad[:, 4:28, 4:28] = imgs - np.array([S_edge(i) for i in imgs])[:, None, None]
    a = recon(P)
    ths = [threading.Thread(target=lambda k=k: [a.insert_slice(pad[i], Transform(mats[i]), 1.0)
                                               for i in range(k, 24, 4)]) for k in range(4)]
    [t.start() for t in ths]
    [t.join() for t in ths]
    b = recon(P)
    for i in range(24):
        b.insert_slice(pad[i], Transform(mats[i]), 1.0)
    ad, an = a.get_state()
    bd, bn = b.get_state()
    assert_state_close(ad, an, bd, bn, tol=1e-5)


def S_edge(img):
    b = np.concatenate([img[0], img[-1], img[:, 0], img[:, -1]]).astype(np.float64)
    return np.float32(b.mean())


AGREE_KEYS = ("reconstruct_norm", "reconstruct_absqual", "reconstruct_absqual_lowres", "reconstruct_qual",
              "reconstruct_weight")


@pytest.mark.parametrize("sym,mode", [("c1", "gauss_2"), ("d2", "trilinear")])
def test_slice_agreement(oracle, sym, mode):
    """determine_slice_agreement (reconstructor.cpp:1212-1381) vs the oracle:
    the five outputs agree to 1e-4 relative.  Enough particles that no voxel
    is reached by the subtracted slice alone: such a voxel keeps only the
    rounding residue of +w then -w, which is summation-order dependent in the
    reference as well (60 particles at P=48 leave 28 of them)."""
    from eman2_amd.reconstructor import EMData
    P, B, n = 32, 24, 600
    ph = S.phantom(10, 7.0, seed=41)
    imgs, mats, _ = S.particles(ph, n, B, seed=42, snr=0.5)
    w = np.ones(n, np.float32)
    r = recon(P, sym, mode)
    r.insert_batch(imgs, list(mats), w)
    slices = [oracle.prep_particle(im, P) for im in imgs]
    v = oracle_volume(oracle, P, sym, mode, slices, mats, w)
    for i, sub in ((0, True), (7, False), (13, True)):
        g = r.determine_slice_agreement(EMData(slices[i], True, {"reconstruct_preproc": 1, "apix_x": 1.5}),
                                        Transform(mats[i]), 1.0, sub)
        o = v.agreement(slices[i], mats[i], 1.0, sub=sub, apix=1.5)
        for k in AGREE_KEYS:
            assert abs(g[k] - o[k]) <= 1e-4 * max(1.0, abs(o[k])), (i, k, g[k], o[k])
    # the volume is unchanged by sub/compare/re-add up to rounding
    gd, gn = r.get_state()
    assert_state_close(gd, gn, v.data, v.norm)


def test_agreement_batch_matches_sequential(oracle):
    P, B, n = 32, 24, 500
    ph = S.phantom(10, 7.0, seed=43)
    imgs, mats, _ = S.particles(ph, n, B, seed=44, snr=0.5)
    w = np.ones(n, np.float32)
    r = recon(P)
    r.insert_batch(imgs, list(mats), w)
    got = r.agreement_batch(imgs[:12], mats[:12], w[:12], sub=True, apix=1.0)
    v = oracle_volume(oracle, P, "c1", "gauss_2", [oracle.prep_particle(im, P) for im in imgs], mats, w)
    for i in range(12):
        o = v.agreement(oracle.prep_particle(imgs[i], P), mats[i], 1.0, sub=True, apix=1.0)
        for j, k in enumerate(AGREE_KEYS):
            assert abs(got[i, j] - o[k]) <= 1e-4 * max(1.0, abs(o[k])), (i, k, got[i, j], o[k])


@pytest.mark.parametrize("sym,mode,P", [("c4", "gauss_2", 48), ("d2", "trilinear", 32), ("c1", "trilinear", 48)])
def test_agreement_batch_symmetry_and_modes(oracle, sym, mode, P):
    """The batched agreement's own-contribution gather under symmetry (the
    other copies' cells read from global memory) and with trilinear weights,
    against the oracle's remove / compare / re-add, sub on and off."""
    B, n = 2 * P // 3, 600
    ph = S.phantom(B // 2 - 2, B / 4.0, seed=47)
    imgs, mats, _ = S.particles(ph, n, B, seed=48, snr=0.5)
    w = np.ones(n, np.float32)
    r = recon(P, sym, mode)
    r.insert_batch(imgs, list(mats), w)
    slices = [oracle.prep_particle(im, P) for im in imgs]
    v = oracle_volume(oracle, P, sym, mode, slices, mats, w)
    for sub in (True, False):
        got = r.agreement_batch(imgs[:10], mats[:10], w[:10], sub=sub, apix=1.0)
        for i in range(10):
            o = v.agreement(slices[i], mats[i], 1.0, sub=sub, apix=1.0)
            for j, k in enumerate(AGREE_KEYS):
                assert abs(got[i, j] - o[k]) <= 1e-4 * max(1.0, abs(o[k])), (sub, i, k, got[i, j], o[k])


def _hip_runtime():
    """The HIP runtime libfourier_hip.so is linked against (loaded with it)."""
    import ctypes

    from eman2_amd import _lib
    _lib.load()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64" in line}
    assert paths, "libamdhip64 not loaded"
    rocm = sorted(q for q in paths if q.startswith("/opt/rocm"))  # the system runtime, not a bundled copy
    hip = ctypes.CDLL((rocm or sorted(paths))[0])
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def test_agreement_batch_device_pointers():
    """Device-resident particles and output (FH_DEVICE) give the host path's
    results: the same kernels, only the staging differs."""
    import ctypes
    P, B, n, m = 32, 24, 200, 40
    ph = S.phantom(10, 7.0, seed=45)
    imgs, mats, dfs = S.particles(ph, n, B, seed=46, snr=0.5, ctf=dict(voltage=300.0, cs=2.7, ampcont=10.0))
    w = np.ones(n, np.float32)
    ctf = dict(voltage=300.0, cs=2.7, ampcont=10.0, apix=1.0)
    r = recon(P)
    r.insert_batch(imgs, list(mats), w, defocus=dfs, ctf=ctf)
    host = r.agreement_batch(imgs[:m], mats[:m], w[:m], defocus=dfs[:m], ctf=ctf, sub=True, apix=1.0)
    hip = _hip_runtime()
    src = np.ascontiguousarray(imgs[:m], np.float32)
    dimg, dout = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dimg), src.nbytes) == 0
    assert hip.hipMalloc(ctypes.byref(dout), m * 5 * 4) == 0
    try:
        assert hip.hipMemcpy(dimg, src.ctypes.data, src.nbytes, 1) == 0  # host to device
        res = r.agreement_batch(dimg.value, mats[:m], w[:m], defocus=dfs[:m], ctf=ctf, sub=True, apix=1.0, box=B,
                                out=dout.value)
        r.sync()
        got = np.zeros((m, 5), np.float32)
        assert hip.hipMemcpy(got.ctypes.data, dout, got.nbytes, 2) == 0  # device to host
        assert res is None
        np.testing.assert_allclose(got, host, rtol=1e-6, atol=1e-7)
        with pytest.raises(RuntimeError, match="device output"):
            r.agreement_batch(dimg.value, mats[:m], w[:m], box=B)
    finally:
        hip.hipFree(dimg)
        hip.hipFree(dout)


@pytest.mark.parametrize("sym,mode,P", [("c1", "gauss_2", 48), ("d7", "trilinear", 48), ("d2", "gauss_2", 96)])
def test_graph_batch_matches_eager(sym, mode, P):
    """The hipGraph form (device-resident particles, one captured graph per
    chunk, 4 chunks) against the eager form on the same device particles,
    three graph runs with eager runs and synchronisations between them;
    every graph run is compared with the eager result.  Pad 96 takes the
    column-first prep (persistent passes) inside the captured graph."""
    import ctypes
    B, n = 2 * P // 3, 200
    ph = S.phantom(12, 10.0, seed=47)
    imgs, mats, dfs = S.particles(ph, n, B, seed=48, snr=0.5, ctf=dict(voltage=300.0, cs=2.7, ampcont=10.0))
    w = np.ones(n, np.float32)
    ctf = dict(voltage=300.0, cs=2.7, ampcont=10.0, apix=1.0)
    hip = _hip_runtime()
    src = np.ascontiguousarray(imgs, np.float32)
    dimg = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dimg), src.nbytes) == 0
    try:
        assert hip.hipMemcpy(dimg, src.ctypes.data, src.nbytes, 1) == 0
        r = recon(P, sym, mode)
        t = r.get_tuning()
        r.set_tuning(**dict(t, chunk=64, graph=0))
        r.setup()
        r.insert_batch(dimg.value, list(mats), w, defocus=dfs, ctf=ctf, box=B)
        ed, en = r.get_state()
        for k in range(3):
            r.set_tuning(**dict(t, chunk=64, graph=1))
            r.setup()
            r.insert_batch(dimg.value, list(mats), w, defocus=dfs, ctf=ctf, box=B)
            gd, gn = r.get_state()
            st = r.get_stats()
            assert st["chunks"] == 4 and st["total_ms"] > 0
            assert_state_close(gd, gn, ed, en, tol=1e-5)
            r.set_tuning(**dict(t, chunk=64, graph=0))  # eager work between the graph runs
            r.setup()
            r.insert_batch(dimg.value, list(mats), w, defocus=dfs, ctf=ctf, box=B)
            r.sync()
    finally:
        hip.hipFree(dimg)


def test_graph_batch_with_2d_transforms_and_ssnr():
    """graph form with the 2-D resampling stage and class_ssnr weights, against
    the eager form on the same device particles (two graph runs, eager work
    between them)"""
    import ctypes
    P, B, n = 48, 32, 160
    ph = S.phantom(12, 10.0, seed=51)
    imgs, mats, _ = S.particles(ph, n, B, seed=52, snr=0.5)
    xfs = _with_2d(mats, 53, True)
    ss = np.random.default_rng(54).uniform(0.0, 2.0, (n, 13)).astype(np.float32)
    hip = _hip_runtime()
    src = np.ascontiguousarray(imgs, np.float32)
    dimg = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dimg), src.nbytes) == 0
    try:
        assert hip.hipMemcpy(dimg, src.ctypes.data, src.nbytes, 1) == 0
        r = recon(P, "d2", "gauss_2", usessnr=True)
        t = r.get_tuning()
        r.set_tuning(**dict(t, chunk=64, graph=0))
        r.setup()
        r.insert_batch(dimg.value, list(xfs), None, box=B, class_ssnr=ss)
        ed, en = r.get_state()
        # the eager device run against host particles (same kernels, staged copies)
        h = recon(P, "d2", "gauss_2", usessnr=True)
        h.insert_batch(imgs, list(xfs), None, class_ssnr=ss)
        hd, hn = h.get_state()
        assert_state_close(ed, en, hd, hn, tol=1e-5)
        for k in range(2):
            r.set_tuning(**dict(t, chunk=64, graph=1))
            r.setup()
            r.insert_batch(dimg.value, list(xfs), None, box=B, class_ssnr=ss)
            gd, gn = r.get_state()
            assert_state_close(gd, gn, ed, en, tol=1e-5)
            r.set_tuning(**dict(t, chunk=64, graph=0))
            r.setup()
            r.insert_batch(dimg.value, list(xfs), None, box=B, class_ssnr=ss)
            r.sync()
    finally:
        hip.hipFree(dimg)


def test_singular_transform_rejected():
    P = 16
    r = recon(P)
    img = np.zeros((P, P), np.float32)
    assert r.insert_slice(img, Transform({"type": "eman", "az": 10, "tx": 2.0}), 1.0) == 0
    with pytest.raises(RuntimeError, match="determinant"):
        r.insert_slice(img, np.zeros((3, 4), np.float32), 1.0)
