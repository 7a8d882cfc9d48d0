"""The CTF_SIGN test of the prep kernel (ctf_sign in prep_kernels.hip):
sign(cos t) taken as the parity of floor(t/pi - 1/2) away from the zeros of
cos, cos itself next to them.  Restated here and checked against numpy's cos
on random phases and on the float values nearest each zero.  (CPU only.)"""
import math

import numpy as np


def parity_sign(t):
    u = t * 0.31830988618379067154 - 0.5
    fl = math.floor(u)
    fr = u - fl
    if 1e-9 < fr < 1.0 - 1e-9:
        return -1.0 if (int(fl) & 1) == 0 else 1.0
    return -1.0 if math.cos(t) < 0 else 1.0


def test_sign_matches_cos_on_random_phases():
    rng = np.random.default_rng(0)
    ts = np.concatenate([rng.uniform(-50, 50, 20000), rng.uniform(-3000, 3000, 20000)]).astype(np.float32)
    for t in ts.astype(np.float64):
        assert parity_sign(t) == (-1.0 if math.cos(t) < 0 else 1.0), t


def test_sign_next_to_zeros_of_cos():
    for k in range(-2000, 2000):
        z = np.float32((k + 0.5) * math.pi)
        for t in (np.nextafter(z, np.float32(-1e9)), z, np.nextafter(z, np.float32(1e9))):
            t = float(t)
            assert parity_sign(t) == (-1.0 if math.cos(t) < 0 else 1.0), (k, t)
