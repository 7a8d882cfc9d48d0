"""GPU parity of the fourier_iter reconstructor ("fourier_iter_hip";
FourierIterReconstructor, libEM/reconstructor.cpp:428-682) against the
oracle: the unseeded Gaussian gridding under c1 and d2, the reference
volume set from a real map on the GPU (do_fft with the phase origin moved to
the corner) or from a complex seed, the seeded insertion, and e2make3dpar's
--iterative loop.  Bars: weights per voxel 1e-4, sums and maps
max|a-b|/max|b| <= 1e-4 with FSC >= 0.9999 (FP32 atomics: summation order)."""
import numpy as np
import pytest

from eman2_amd import make3d
from eman2_amd import synthetic as S
from eman2_amd.metrics import fsc, max_rel_err, weight_rel_err

pytestmark = pytest.mark.gpu
CTF = dict(voltage=300.0, cs=2.7, ampcont=10.0, apix=1.0)


def recon(P, sym="c1"):
    from eman2_amd import Reconstructors
    print("fourier_iter P=%d sym=%s" % (P, sym), flush=True)
    return Reconstructors.get("fourier_iter_hip", dict(size=(P, P, P), sym=sym))


def inputs(n, B=32, seed=81):
    ph = S.phantom(12, B * 0.3, seed=seed)
    imgs, mats, dfs = S.particles(ph, n, B, seed=seed + 1, snr=0.5, ctf=dict(CTF, bfactor=80.0))
    return ph, imgs, mats, dfs


def oracle_slices(O, imgs, dfs, P):
    return [O.prep_particle(imgs[i], P, dict(CTF, defocus=float(dfs[i]))) for i in range(len(imgs))]


def close(gd, gn, od, on):
    e1, e2 = weight_rel_err(gn, on), max_rel_err(gd, od)
    print("weights %.2e sums %.2e" % (e1, e2), flush=True)
    assert e1 <= 1e-4 and e2 <= 1e-4


@pytest.mark.parametrize("sym", ["c1", "d2"])
def test_iter_unseeded(oracle, sym):
    P, n = 48, 200
    _, imgs, mats, dfs = inputs(n)
    r = recon(P, sym)
    r.setup()
    r.insert_batch(imgs, mats, np.ones(n, np.float32), defocus=dfs, ctf=CTF)
    gd, gn = r.get_state()
    v = oracle.Volume(P, sym, "fourier_iter")
    for s, m in zip(oracle_slices(oracle, imgs, dfs, P), mats):
        v.insert(s, m, 1.0)
    close(gd, gn, v.data, v.norm)
    g = r.finish(True, return_numpy=True)
    o = v.finish(True)
    assert max_rel_err(g, o) <= 1e-4 and fsc(g, o).min() >= 0.9999


def test_iter_seeded_from_map_and_complex(oracle):
    P, n = 48, 150
    ph, imgs, mats, dfs = inputs(n, seed=91)
    seed_map = (S.phantom_volume(ph, P) * 40.0).astype(np.float32)
    w = np.ones(n, np.float32)
    slices = oracle_slices(oracle, imgs, dfs, P)
    v = oracle.Volume(P, "c1", "fourier_iter")
    v.setup_iter_ref(seed_map)
    for s, m in zip(slices, mats):
        v.insert(s, m, 1.0)
    r = recon(P)
    r.setup_seed(seed_map, 1.0)                  # the GPU FFT of the map
    r.insert_batch(imgs, mats, w, defocus=dfs, ctf=CTF)
    close(*r.get_state(), v.data, v.norm)
    # a complex seed is taken as the reference as it is
    r2 = recon(P)
    r2.setup_seed(v.ref, 1.0)
    r2.insert_batch(imgs, mats, w, defocus=dfs, ctf=CTF)
    close(*r2.get_state(), v.data, v.norm)
    # setup() drops the reference again: the unseeded insertion
    r2.setup()
    r2.insert_batch(imgs[:20], mats[:20], w[:20], defocus=dfs[:20], ctf=CTF)
    u = oracle.Volume(P, "c1", "fourier_iter")
    for s, m in zip(slices[:20], mats[:20]):
        u.insert(s, m, 1.0)
    close(*r2.get_state(), u.data, u.norm)


class OracleIter:
    """the oracle behind the calls reconstruct_iterative makes"""

    def __init__(self, O, P):
        self.O, self.P = O, P
        self.v = O.Volume(P, "c1", "fourier_iter")

    def setup(self):
        self.v = self.O.Volume(self.P, "c1", "fourier_iter")

    def setup_seed(self, seed, w):
        self.setup()
        self.v.setup_iter_ref(seed)

    def insert_batch(self, imgs, xforms, weights, defocus=None, ctf=None):
        for i in range(len(imgs)):
            self.v.insert(self.O.prep_particle(imgs[i], self.P, dict(ctf, defocus=float(defocus[i]))), xforms[i],
                          float(weights[i]))

    def finish(self, doift=True):
        return self.v.finish(True)


def test_iterative_loop(oracle):
    """e2make3dpar --iterative: passes seeded with the previous, compressed map"""
    P, n = 48, 150
    _, imgs, mats, dfs = inputs(n, seed=101)
    w = np.ones(n, np.float32)
    g = make3d.reconstruct_iterative(recon(P), imgs, mats, w, niter=3, defocus=dfs, ctf=CTF)
    o = make3d.reconstruct_iterative(OracleIter(oracle, P), imgs, mats, w, niter=3, defocus=dfs, ctf=CTF)
    e, f = max_rel_err(g, o), fsc(g, o).min()
    print("iterative: max_rel_err %.2e min FSC %.6f" % (e, f), flush=True)
    assert e <= 1e-4 and f >= 0.9999


def test_iter_agreement_is_a_no_op():
    P, n = 32, 40
    _, imgs, mats, dfs = inputs(n, B=24, seed=111)
    r = recon(P)
    r.setup()
    r.insert_batch(imgs, mats, np.ones(n, np.float32), defocus=dfs, ctf=CTF)
    before = r.get_state()
    assert r.determine_slice_agreement(np.zeros((P, P), np.float32), mats[0], 1.0, True) == {}
    out = r.agreement_batch(imgs[:4], mats[:4], np.ones(4, np.float32), defocus=dfs[:4], ctf=CTF)
    assert np.all(out == 0)
    after = r.get_state()
    assert all(np.array_equal(a, b) for a, b in zip(before, after))
