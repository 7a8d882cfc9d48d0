"""CPU replay of ctf_sign (prep_kernels.hip): the single-precision test of
floor(t / pi - 1/2) decides the CTF_SIGN of libEM/ctf.cpp:1076-1090 exactly
as the double-precision one wherever it decides (|u| < 4096, fraction more
than 1e-3 from an integer), over every Fourier pixel of pad 384 and 600
slices at 120 defocus values and several microscopes.  The fma is replayed
in extended precision (its 48-bit product and the 1/2 fit the 64-bit
significand), so the float32 result is its single rounding.  (CPU only.)"""
import math

import numpy as np
import pytest

f32 = np.float32


def consts(voltage, cs, ampcont, apix, ny):
    lam = f32(12.2639) / np.sqrt(f32(voltage * 1000.0) + f32(0.97845) * f32(voltage) * f32(voltage), dtype=f32)
    phase = f32(math.asin(ampcont / 100.0))
    ds = f32(1.0) / (f32(apix) * f32(ny))
    g1 = f32(math.pi / 2.0 * cs * 1.0e7 * float(lam) ** 3)
    g2 = f32(math.pi * float(lam) * 10000.0)
    acac = f32(math.pi / 2.0 - float(phase))
    return g1, g2, acac, ds


@pytest.mark.parametrize("P,voltage,cs,ampcont,apix", [(384, 300.0, 2.7, 10.0, 1.0), (600, 300.0, 2.7, 10.0, 0.8),
                                                       (384, 200.0, 0.01, 0.0, 1.5), (600, 300.0, 2.7, 50.0, 2.5)])
def test_single_precision_decisions_match_double(P, voltage, cs, ampcont, apix):
    g1, g2, acac, ds = consts(voltage, cs, ampcont, apix, P)
    H = P // 2
    kx, ky = np.meshgrid(np.arange(H + 1), np.arange(H + 1))
    n = (kx * kx + ky * ky).ravel()
    hyp = np.sqrt(n.astype(np.float64)).astype(f32)  # correctly rounded (hypot_fast)
    decided = 0
    for df in np.linspace(0.2, 6.0, 120).astype(f32):
        s = (hyp * ds).astype(f32)
        a = ((((-g1) * s).astype(f32) * s).astype(f32) * s).astype(f32) * s
        b = (((g2 * df).astype(f32) * s).astype(f32) * s).astype(f32)
        gam = (a.astype(f32) + b).astype(f32)
        tf = (gam - acac).astype(f32)
        # double path: u = t * (1/pi) - 1/2, two roundings
        t = tf.astype(np.float64)
        u = t * 0.31830988618379067154 - 0.5
        fl = np.floor(u)
        fr = u - fl
        sure = (fr > 1e-9) & (fr < 1.0 - 1e-9)
        # float path: fma(tf, fl(1/pi), -1/2) rounded once
        uf = (tf.astype(np.longdouble) * np.longdouble(f32(0.318309886)) - np.longdouble(0.5)).astype(f32)
        flf = np.floor(uf)
        frf = (uf - flf).astype(f32)
        fast = (np.abs(uf) < f32(4096.0)) & (frf > f32(1e-3)) & (frf < f32(1.0) - f32(1e-3))
        assert np.all(sure[fast])  # the fast path only decides where the double path's parity does
        assert np.array_equal(flf[fast].astype(np.int64) & 1, fl[fast].astype(np.int64) & 1)
        decided += int(fast.sum())
    assert decided > 0.99 * 120 * n.size
