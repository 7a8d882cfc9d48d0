"""Writes the committed golden fixtures under tests/golden/.

  tables.json      checksums of the fast_exp table (libEM/util.cpp:792-794)
                   and the hypot_fast_int table (libEM/util.cpp:764-785),
                   regenerated from their closed forms, plus the slice pixel
                   counts N_pix(P) kept by the radius cut of
                   libEM/reconstructor.cpp:1168-1181, and the reference's own
                   FFT known answers (rt/pyem/test_emdata.py:722-760).
  recon_p16.npz    a tiny reconstruction (P=16, c1 and d2, gauss_2 and
                   trilinear) computed by the oracle: inputs and outputs.
                   GPU tests compare against it; the oracle test checks the
                   oracle still reproduces it.

Run:  python tests/golden/make_golden.py
"""
import hashlib
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def closed_form_tables():
    fe = np.array([np.float32(math.exp(-i / 50.0)) for i in range(1000)], np.float32)
    dim = 301
    hy = np.zeros((dim, dim), np.int16)
    for y in range(dim):
        for x in range(dim):
            hy[y, x] = int(math.floor(math.hypot(x, y) + 0.5))
    return fe, hy


def npix(P):
    h = P // 2
    n = 0
    for y in range(-h, h):
        for x in range(0, h + 1):
            if int(math.floor(math.hypot(x, y) + 0.5)) <= h:
                n += 1
    return n


def main():
    fe, hy = closed_form_tables()
    tables = {
        "fast_exp_sha256": hashlib.sha256(fe.tobytes()).hexdigest(),
        "hypot_fast_int_301_sha256": hashlib.sha256(hy.tobytes()).hexdigest(),
        "npix": {str(P): npix(P) for P in (16, 32, 96, 192, 288, 384, 480, 600)},
        # rt/pyem/test_emdata.py:722-760 (assertAlmostEqual to 2 places)
        "fft_known_answers": {
            "1d_1..5": [15, 0, -2.5, 3.44, -2.5, 0.81],
            "1d_1..4": [10, 0, -2, 2, -2, 0],
            "2x2_1..4": [[10, 0, -2, 0], [-4, 0, 0, 0]],
        },
    }
    with open(os.path.join(HERE, "tables.json"), "w") as f:
        json.dump(tables, f, indent=1, sort_keys=True)

    from oracle import oracle as O
    from eman2_amd import synthetic as S
    P = 16
    ph = S.phantom(6, 4.0, seed=11, sigma=(1.0, 2.0))
    imgs, mats, _ = S.particles(ph, 6, P, seed=12, snr=1.0)
    weights = np.array([1.0, 0.5, 2.0, 1.0, 1.0, 0.75], np.float32)
    out = {"imgs": imgs, "mats": mats, "weights": weights}
    slices = np.stack([O.prep_particle(im, P) for im in imgs])
    out["slices"] = slices
    for sym in ("c1", "d2"):
        for mode in ("gauss_2", "trilinear"):
            v = O.Volume(P, sym, mode)
            for s, m, w in zip(slices, mats, weights):
                v.insert(s, m, w)
            key = "%s_%s" % (sym, mode)
            out[key + "_data"] = v.data.copy()
            out[key + "_norm"] = v.norm.copy()
            out[key + "_map"] = v.finish(True)
    np.savez_compressed(os.path.join(HERE, "recon_p16.npz"), **out)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
