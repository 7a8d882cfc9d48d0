"""The e2make3d.py reconstruct() loop (programs/e2make3d.py:385-539) over the
HIP reconstructor against the same loop over the oracle: per-particle norms
and qualities within 1e-4, final maps max|a-b|/max|b| <= 1e-4 with FSC >=
0.9999; the quality cutoff leaves out the requested share."""
import numpy as np
import pytest

from eman2_amd import make3d
from eman2_amd import synthetic as S
from eman2_amd.metrics import fsc, max_rel_err

pytestmark = pytest.mark.gpu


class OracleRecon:
    """the oracle behind the reconstructor calls make3d.reconstruct makes"""

    def __init__(self, O, P, sym="c1", mode="gauss_2"):
        self.O, self.P, self.sym, self.mode = O, P, sym, mode

    def setup(self):
        self.v = self.O.Volume(self.P, self.sym, self.mode)

    def insert_batch(self, imgs, xforms, weights, defocus=None, ctf=None):
        for im, m, w in zip(imgs, xforms, weights):
            self.v.insert(self.O.prep_particle(im, self.P), m, float(w))

    def agreement_batch(self, imgs, xforms, weights, defocus=None, ctf=None, sub=True, apix=1.0):
        out = np.zeros((len(imgs), 5), np.float32)
        for i, (im, m, w) in enumerate(zip(imgs, xforms, weights)):
            r = self.v.agreement(self.O.prep_particle(im, self.P), m, float(w), sub=sub, apix=apix)
            out[i] = [r[k] for k in make3d.AGREE_KEYS]
        return out

    def finish(self, doift=True):
        return self.v.finish(True)


def _inputs(n=300, P=32, B=24):
    ph = S.phantom(10, 7.0, seed=61)
    imgs, mats, _ = S.particles(ph, n, B, seed=62, snr=0.5)
    w = np.ones(n, np.float32)
    w[::17] = 0.0  # some particles with weight 0
    return P, imgs, mats, w


def test_make3d_loop_matches_oracle(oracle):
    from eman2_amd import Reconstructors
    P, imgs, mats, w = _inputs()
    print("make3d P=%d n=%d" % (P, len(imgs)), flush=True)
    r = Reconstructors.get("fourier_hip", dict(size=(P, P, P), sym="c1", mode="gauss_2"))
    gmap, ginfo = make3d.reconstruct(r, imgs, mats, w, niter=2, keep=1.0)
    omap, oinfo = make3d.reconstruct(OracleRecon(oracle, P), imgs, mats, w, niter=2, keep=1.0)
    g = gmap.numpy() if hasattr(gmap, "numpy") else np.asarray(gmap)
    assert max_rel_err(g, omap) <= 1e-4
    assert fsc(g, omap).min() >= 0.9999
    np.testing.assert_allclose(ginfo["norm"], oinfo["norm"], rtol=1e-4)
    np.testing.assert_allclose(ginfo["reconstruct_absqual"], oinfo["reconstruct_absqual"], rtol=1e-4, atol=1e-6)
    assert ginfo["ptcl_repr"] == oinfo["ptcl_repr"] == float(w.sum())


def test_make3d_quality_cutoff():
    from eman2_amd import Reconstructors
    P, imgs, mats, w = _inputs()
    r = Reconstructors.get("fourier_hip", dict(size=(P, P, P), sym="c1", mode="gauss_2"))
    _, info = make3d.reconstruct(r, imgs, mats, w, niter=3, keep=0.8)
    n = len(imgs)
    nz = int((w != 0).sum())
    excl = set(info["threed_excl_ptcl_idxs"])
    assert set(np.nonzero(w == 0)[0]) <= excl
    # qcutoff = the int(0.8 * len) -th largest relative quality: the rest are left out
    assert len(excl) - (n - nz) == nz - int(0.8 * nz)
    assert len(info["threed_ptcl_idxs"]) + len(excl) == n


def test_make3d_from_mrc_stack(tmp_path):
    """the same loop fed from an MRC particle stack, the map written back"""
    from eman2_amd import Reconstructors, mrc
    P, imgs, mats, w = _inputs(n=120)
    stack = tmp_path / "ptcls.mrcs"
    mrc.write(str(stack), imgs, apix=1.0)
    r = Reconstructors.get("fourier_hip", dict(size=(P, P, P), sym="c1", mode="gauss_2"))
    out_path = tmp_path / "map.mrc"
    out, info = make3d.reconstruct_mrc(r, str(stack), mats, w, out_path=str(out_path), niter=1)
    r2 = Reconstructors.get("fourier_hip", dict(size=(P, P, P), sym="c1", mode="gauss_2"))
    ref, _ = make3d.reconstruct(r2, imgs, mats, w, niter=1)
    vol, h = mrc.read(str(out_path))
    g = ref.numpy() if hasattr(ref, "numpy") else np.asarray(ref)
    assert h["nx"] == P and h["apix"] == 1.0
    assert max_rel_err(np.asarray(vol), g) <= 1e-4


def test_spa_make3d_task_merge_equals_one_reconstructor():
    """e2spa_make3d's split into tasks, each finished to its normalised
    volume and normout, merged by summing volume x normout and normout, equals
    one reconstructor over all particles (the weight fix-ups are linear)"""
    from eman2_amd import Reconstructors
    P, B, n = 48, 32, 400
    ph = S.phantom(12, 9.0, seed=71)
    ctf = dict(voltage=300.0, cs=2.7, ampcont=10.0, apix=1.0)
    imgs, mats, dfs = S.particles(ph, n, B, seed=72, snr=0.5, ctf=dict(ctf, bfactor=80.0))
    w = np.ones(n, np.float32)
    w[::13] = 0.0
    mk = lambda: Reconstructors.get("fourier_hip", dict(size=(P, P, P), sym="d2", mode="gauss_2"))  # noqa: E731
    merged = make3d.spa_make3d(mk, imgs, mats, w, ntasks=3, defocus=dfs, ctf=ctf)
    one = mk()
    one.setup()
    one.insert_batch(imgs, mats, w, defocus=dfs, ctf=ctf)
    ref = one.finish(True, return_numpy=True)
    e, f = max_rel_err(merged, ref), fsc(merged, ref).min()
    print("spa merge: max_rel_err %.2e min FSC %.6f" % (e, f), flush=True)
    assert e <= 1e-4 and f >= 0.9999


@pytest.mark.parametrize("mode", ["gauss_2", "gauss_var"])
def test_fillangle_insertion_matches_oracle(oracle, mode):
    """e2make3dpar.py --fillangle (:737-747): each particle, preprocessed
    with its own Transform (here with a translation), inserted at the den x
    den (alt, az) grid of fill_orientations with phi 0 and the Gaussian
    factor -- against the oracle's prep with the particle's Transform and
    insertion at each fill rotation"""
    from eman2_amd import Reconstructors
    from eman2_amd.transform import euler_eman_matrix
    P, B, n = 48, 32, 10
    ph = S.phantom(10, 10.0, seed=71)
    imgs, mats, _ = S.particles(ph, n, B, seed=72, snr=1.0)
    rng = np.random.default_rng(73)
    xf = np.array(mats, np.float32).reshape(n, 3, 4).copy()
    xf[:, :2, 3] = rng.uniform(-2.0, 2.0, (n, 2))
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    fa = 6.0   # P = 48: step atan2(1, 24) = 2.39 deg, den = 2
    print("fillangle %s P=%d n=%d" % (mode, P, n), flush=True)
    r = Reconstructors.get("fourier_hip", dict(size=(P, P, P), sym="c1", mode=mode))
    r.setup()
    make3d.insert_with_fill(r, imgs, xf, w, fillangle=fa)
    gd, gn = r.get_state()
    v = oracle.Volume(P, "c1", mode)
    for i in range(n):
        s = oracle.prep_particle(imgs[i], P, xform=xf[i])
        for mk, wm in make3d.fill_orientations(xf[i], fa, P):
            rot = mk.copy()
            rot[:, 3] = 0.0   # the insertion takes the rotation; the translation went to the prep
            v.insert(s, rot, float(np.float32(w[i] * wm)))
    assert len(make3d.fill_orientations(xf[0], fa, P)) == 4
    from eman2_amd.metrics import weight_rel_err
    assert weight_rel_err(gn, v.norm) <= 1e-4
    assert max_rel_err(gd, v.data) <= 1e-4
