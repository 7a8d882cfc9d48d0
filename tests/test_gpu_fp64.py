"""Error of the GPU volume against an FP64 re-run of the same insertion, per
workload shape: config 1 at full size (2000 particles, 64 px box, pad 96,
SNR 0.1, no CTF), and the box / pad / symmetry / CTF of the other configs
with fewer particles so that the FP64 oracle finishes in seconds.  The FP64
run accumulates in double (oracle or_insert_slice_work_f64) with the same
per-contribution arithmetic as the reference.  Reported (printed) per shape:
max|a-b|/max|b| of the complex sums, per-voxel relative weight error, the
finished maps' max relative error and minimum FSC to Nyquist.
Bars: sums and weights <= 1e-4, maps <= 1e-4 and FSC >= 0.9999."""
import numpy as np
import pytest

from eman2_amd import synthetic as S
from eman2_amd.metrics import fsc, max_rel_err, weight_rel_err

pytestmark = pytest.mark.gpu
CTF = dict(voltage=300.0, cs=2.7, ampcont=10.0, bfactor=80.0)

# (name, box, pad, apix, particles, sym, mode, ctf, snr, n_gauss)
SHAPES = [
    ("config 1 (full)", 64, 96, 1.5, 2000, "c1", "gauss_2", False, 0.1, 50),
    ("config 2 shape", 256, 384, 1.0, 200, "c1", "gauss_2", True, 0.05, 200),
    ("config 3 shape, gauss_2", 192, 288, 1.0, 60, "d7", "gauss_2", True, 0.05, 60),
    ("config 3 shape, trilinear", 192, 288, 1.0, 60, "d7", "trilinear", True, 0.05, 60),
    ("config 5 shape", 400, 600, 0.8, 40, "c1", "gauss_2", True, 0.05, 100),
    ("config 1 shape, gauss_var", 64, 96, 1.5, 300, "c1", "gauss_var", False, 0.1, 50),
]


@pytest.mark.parametrize("shape", SHAPES, ids=[s[0] for s in SHAPES])
def test_error_vs_fp64_rerun(oracle, shape):
    from eman2_amd import Reconstructors
    name, B, P, apix, n, sym, mode, use_ctf, snr, ng = shape
    print("fp64 re-run:", name, flush=True)
    ph = S.phantom(ng, 0.375 * B, seed=7)
    imgs, mats, dfs = S.particles(ph, n, B, seed=8, apix=apix, snr=snr, ctf=dict(CTF) if use_ctf else None)
    w = np.ones(n, np.float32)
    r = Reconstructors.get("fourier_hip", {"size": (P, P, P), "sym": sym, "mode": mode})
    r.setup()
    ctf = dict(CTF, apix=apix) if use_ctf else None
    r.insert_batch(imgs, list(mats), w, defocus=dfs if use_ctf else None, ctf=ctf)
    gd, gn = r.get_state()
    gmap = r.finish(True, return_numpy=True)

    v64 = oracle.Volume(P, sym, mode, fp64=True)
    for i in range(n):
        c = dict(CTF, apix=apix, defocus=float(dfs[i])) if use_ctf else None
        v64.insert(oracle.prep_particle(imgs[i], P, c), mats[i], 1.0)
    v32 = oracle.Volume(P, sym, mode)
    v32.data = v64.data.astype(np.float32)
    v32.norm = v64.norm.astype(np.float32)
    omap = v32.finish(True)

    e_sum = max_rel_err(gd, v64.data)
    e_w = weight_rel_err(gn, v64.norm)
    e_map = max_rel_err(gmap, omap)
    f = fsc(gmap, omap).min()
    print("fp64 re-run: %-28s sums %.2e  weights %.2e  map %.2e  FSC min %.6f" % (name, e_sum, e_w, e_map, f),
          flush=True)
    assert e_sum <= 1e-4 and e_w <= 1e-4
    assert e_map <= 1e-4 and f >= 0.9999
