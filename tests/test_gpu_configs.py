"""GPU parity at the shapes of the serving mix and of the two-pass refine flow.

* Pads 192 and 480 (boxes 128 and 320 of the serving mix): the compile-time
  FFT plans of those pads in the row and column passes, batched insertion
  under c4 and d2 with CTF phase flip, and finish() with its 3-D inverse.
* Batched slice agreement on volumes wider than one 32-pixel x tile of the
  compare kernel (P = 128: three tiles; P = 600: ten), with and without the
  particle's own contribution removed, then the refine flow's re-insertion of
  every particle scaled by its reconstruct_norm and the finished map (pad 600,
  box 400, 0.8 A/px).

Bars as in the other parity tests: slices max|a-b| <= 2e-5 max|b|; weights
per voxel <= 1e-4; sums and maps max|a-b|/max|b| <= 1e-4 with FSC >= 0.9999;
agreement outputs 1e-4 relative.  The oracle's 3-D inverse FFT is a textbook
transform in double that equals numpy's (pinned on the CPU); at these sizes
the maps are compared against numpy's irfftn of the oracle's normalised
volume, which takes seconds instead of the textbook transform's minute.
"""
import numpy as np
import pytest

from eman2_amd import synthetic as S
from eman2_amd.metrics import fsc, max_rel_err, weight_rel_err

pytestmark = pytest.mark.gpu
CTF_GEN = dict(voltage=300.0, cs=2.7, ampcont=10.0, bfactor=80.0)
KEYS = ("reconstruct_norm", "reconstruct_absqual", "reconstruct_absqual_lowres", "reconstruct_qual",
        "reconstruct_weight")


def recon(P, sym="c1", mode="gauss_2", **kw):
    from eman2_amd import Reconstructors
    print("recon P=%d sym=%s mode=%s %s" % (P, sym, mode, kw), flush=True)
    r = Reconstructors.get("fourier_hip", dict(size=(P, P, P), sym=sym, mode=mode, **kw))
    r.setup()
    return r


def cplx(a):
    return a[..., 0::2].astype(np.float64) + 1j * a[..., 1::2].astype(np.float64)


def oracle_map(v, P):
    """the oracle's normalised volume, inverse-transformed by numpy (c2r, 1/P^3,
    phase origin to the centre)"""
    c = cplx(v.finish(False))
    return np.fft.fftshift(np.fft.irfftn(c, s=(P, P, P), axes=(0, 1, 2))).astype(np.float32)


def assert_map_close(a, b, tol=1e-4):
    e, f = max_rel_err(a, b), fsc(a, b).min()
    print("map max_rel_err %.2e  min FSC %.6f" % (e, f), flush=True)
    assert e <= tol and f >= 0.9999, (e, f)


def inputs(B, n, seed, apix, nph):
    ph = S.phantom(nph, B * 0.35, seed=seed)
    imgs, mats, dfs = S.particles(ph, n, B, seed=seed + 1, apix=apix, snr=0.1, ctf=CTF_GEN)
    return imgs, mats, dfs, dict(voltage=300.0, cs=2.7, ampcont=10.0, apix=apix)


@pytest.mark.parametrize("B,P,sym", [(128, 192, "c4"), (320, 480, "d2")])
def test_serving_pads_prep_insert_finish(oracle, B, P, sym):
    n = 24
    imgs, mats, dfs, ctf = inputs(B, n, P, 1.0, 60)
    w = np.ones(n, np.float32)
    r = recon(P, sym)
    g = r.prep_batch(imgs, defocus=dfs, ctf=ctf)
    slices = [oracle.prep_particle(imgs[i], P, dict(ctf, defocus=float(dfs[i]))) for i in range(n)]
    for i in range(n):
        assert max_rel_err(g[i], slices[i]) <= 2e-5, (i, max_rel_err(g[i], slices[i]))
    r.insert_batch(imgs, list(mats), w, defocus=dfs, ctf=ctf)
    gd, gn = r.get_state()
    v = oracle.Volume(P, sym, "gauss_2")
    for s, m in zip(slices, mats):
        v.insert(s, m, 1.0)
    assert weight_rel_err(gn, v.norm) <= 1e-4
    assert max_rel_err(gd, v.data) <= 1e-4
    assert_map_close(r.finish(True, return_numpy=True), oracle_map(v, P))


def _seeded_pair(oracle, P, seed):
    """a GPU reconstructor and an oracle volume holding the same dense base:
    every cell weighted, so that removing a particle's own slice leaves no
    cell with only the rounding residue of +w - w"""
    rng = np.random.default_rng(seed)
    data = (0.05 * rng.standard_normal((P, P, P + 2))).astype(np.float32)
    norm = rng.uniform(0.5, 1.5, (P, P, P // 2 + 1)).astype(np.float32)
    r = recon(P)
    r.setup_seedandweights(data, norm)
    v = oracle.Volume(P, "c1", "gauss_2")
    v.data[:] = data
    v.norm[:] = norm
    return r, v, (data, norm)


@pytest.mark.parametrize("B,P,n,apix", [(96, 128, 24, 1.0), (400, 600, 16, 0.8)])
def test_agreement_batch_across_x_tiles(oracle, B, P, n, apix):
    imgs, mats, dfs, ctf = inputs(B, n, 7 * P, apix, 40)
    w = np.ones(n, np.float32)
    r, v, _ = _seeded_pair(oracle, P, P)
    r.insert_batch(imgs, list(mats), w, defocus=dfs, ctf=ctf)
    slices = [oracle.prep_particle(imgs[i], P, dict(ctf, defocus=float(dfs[i]))) for i in range(n)]
    for s, m in zip(slices, mats):
        v.insert(s, m, 1.0)
    for sub in (True, False):
        got = r.agreement_batch(imgs, mats, w, defocus=dfs, ctf=ctf, sub=sub, apix=apix)
        for i in range(n):
            o = v.agreement(slices[i], mats[i], 1.0, sub=sub, apix=apix)
            for j, k in enumerate(KEYS):
                assert abs(got[i, j] - o[k]) <= 1e-4 * max(1.0, abs(o[k])), (sub, i, k, got[i, j], o[k])


def test_refine_flow_pad600_reinsert_and_finish(oracle):
    """e2make3d.py:447-491 on a seeded pad-600 volume: insert, batched agreement
    with each particle's own slice removed, then the volume rebuilt from the
    base with every particle scaled by its reconstruct_norm, and finished"""
    B, P, n, apix = 400, 600, 16, 0.8
    imgs, mats, dfs, ctf = inputs(B, n, 5, apix, 60)
    w = np.ones(n, np.float32)
    r, v, base = _seeded_pair(oracle, P, 77)
    r.insert_batch(imgs, list(mats), w, defocus=dfs, ctf=ctf)
    slices = [oracle.prep_particle(imgs[i], P, dict(ctf, defocus=float(dfs[i]))) for i in range(n)]
    for s, m in zip(slices, mats):
        v.insert(s, m, 1.0)
    got = r.agreement_batch(imgs, mats, w, defocus=dfs, ctf=ctf, sub=True, apix=apix)
    onorm = np.array([v.agreement(slices[i], mats[i], 1.0, sub=True, apix=apix)["reconstruct_norm"]
                      for i in range(n)], np.float32)
    np.testing.assert_allclose(got[:, 0], onorm, rtol=1e-4)
    # second pass: the base again, particles scaled by their norm
    r.setup_seedandweights(*base)
    scaled = imgs * onorm[:, None, None]
    r.insert_batch(scaled, list(mats), w, defocus=dfs, ctf=ctf)
    v.data[:] = base[0]
    v.norm[:] = base[1]
    for i in range(n):
        v.insert(oracle.prep_particle(scaled[i], P, dict(ctf, defocus=float(dfs[i]))), mats[i], 1.0)
    gd, gn = r.get_state()
    assert weight_rel_err(gn, v.norm) <= 1e-4
    assert max_rel_err(gd, v.data) <= 1e-4
    assert_map_close(r.finish(True, return_numpy=True), oracle_map(v, P))


def _xforms_with_2d(n, seed, scale=False):
    """random orientations with translations (and optionally a scale and an
    x mirror): the 2-D part preprocess_slice resamples with"""
    from eman2_amd.transform import euler_eman_matrix
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        m = euler_eman_matrix(*(rng.random(3) * (360, 180, 360))).astype(np.float32)
        if scale and i % 3 == 1:
            m[:, :3] *= np.float32(1.05)
        if scale and i % 3 == 2:
            m[0, :] *= -1
        m[:2, 3] = rng.uniform(-6.0, 6.0, 2)
        out.append(m)
    return np.stack(out)


@pytest.mark.parametrize("B,P,use_dfdf", [(64, 96, False), (64, 96, True), (256, 384, False), (128, 192, True)])
def test_ctf_flip_then_2d_transform_batched(oracle, B, P, use_dfdf):
    """CTF phase flip together with translations / scale / mirror in the
    batched path, in the reference's order: the padded particle is flipped
    (FFT, CTF_SIGN -- times the sign at defocus + dfdf --, inverse FFT) and
    preprocess_slice then resamples it (e2make3dpar.py:698-716)"""
    n = 12
    rng = np.random.default_rng(B + P)
    imgs = (rng.standard_normal((n, B, B)) + 1.0).astype(np.float32)
    dfs = rng.uniform(0.8, 2.5, n).astype(np.float32)
    mats = _xforms_with_2d(n, P, scale=True)
    ctf = dict(voltage=300.0, cs=2.7, ampcont=10.0, apix=1.0)
    dd = rng.uniform(-0.4, 0.4, n).astype(np.float32) if use_dfdf else None
    if use_dfdf:
        dd[::4] = 0.0            # a zero offset: the two signs cancel
        ctf["dfdf"] = dd
    w = np.ones(n, np.float32)
    r = recon(P, "d2")
    r.insert_batch(imgs, mats, w, defocus=dfs, ctf=ctf)
    gd, gn = r.get_state()
    v = oracle.Volume(P, "d2", "gauss_2")
    oc = dict(voltage=300.0, cs=2.7, ampcont=10.0, apix=1.0)
    for i in range(n):
        s = oracle.prep_particle(imgs[i], P, dict(oc, defocus=float(dfs[i])), xform=mats[i],
                                 dfdf=None if dd is None else float(dd[i]))
        v.insert(s, mats[i], 1.0)
    assert weight_rel_err(gn, v.norm) <= 1e-4
    e = max_rel_err(gd, v.data)
    print("sums max_rel_err %.2e" % e, flush=True)
    assert e <= 1e-4
    # the batched agreement takes the same prep
    got = r.agreement_batch(imgs, mats, w, defocus=dfs, ctf=ctf, sub=False, apix=1.0)
    for i in range(0, n, 3):
        s = oracle.prep_particle(imgs[i], P, dict(oc, defocus=float(dfs[i])), xform=mats[i],
                                 dfdf=None if dd is None else float(dd[i]))
        o = v.agreement(s, mats[i], 1.0, sub=False, apix=1.0)
        for j, k in enumerate(KEYS):
            assert abs(got[i, j] - o[k]) <= 1e-4 * max(1.0, abs(o[k])), (i, k, got[i, j], o[k])


def test_prep_batch_dfdf(oracle):
    """the two-sign flip alone (no 2-D part): prep_batch slices against the oracle"""
    B, P, n = 100, 150, 8
    rng = np.random.default_rng(9)
    imgs = (rng.standard_normal((n, B, B)) + 2.0).astype(np.float32)
    dfs = rng.uniform(0.8, 2.5, n).astype(np.float32)
    dd = rng.uniform(-0.5, 0.5, n).astype(np.float32)
    ctf = dict(voltage=200.0, cs=2.0, ampcont=7.0, apix=1.3)
    g = recon(P).prep_batch(imgs, defocus=dfs, ctf=dict(ctf, dfdf=dd))
    for i in range(n):
        o = oracle.prep_particle(imgs[i], P, dict(ctf, defocus=float(dfs[i])), dfdf=float(dd[i]))
        assert max_rel_err(g[i], o) <= 2e-5, (i, max_rel_err(g[i], o))
