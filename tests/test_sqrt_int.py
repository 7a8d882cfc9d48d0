"""CPU replay of sqrt_int_rn (fh_device.hpp), the correctly rounded sqrt of
an integer n < 2^24 (Util::hypot_fast's rounded hypot of the reference,
libEM/util.cpp:764-785): from any start within one ulp of sqrt(n), as the
hardware sqrt gives, the two fma residual tests land on the correctly
rounded float, for every n.  The fma residual n - a*b is exact in float64
here (a, b have 24-bit significands), and only its sign is used.  (CPU only.)"""
import numpy as np


def corrected(x, s):
    sd = np.nextafter(s, np.float32(-np.inf))
    su = np.nextafter(s, np.float32(np.inf))
    xd, sdd, sud, s64 = x.astype(np.float64), sd.astype(np.float64), su.astype(np.float64), s.astype(np.float64)
    r = np.where(xd - sdd * s64 <= 0, sd, s)
    return np.where(xd - sud * s64 > 0, su, r)


def test_every_integer_below_2_24():
    for lo in range(0, 1 << 24, 1 << 21):
        n = np.arange(max(lo, 1), lo + (1 << 21), dtype=np.int64)
        x = n.astype(np.float32)
        want = np.sqrt(n.astype(np.float64)).astype(np.float32)
        assert np.array_equal(np.sqrt(x), want)  # float32 sqrt is correctly rounded
        for s in (want, np.nextafter(want, np.float32(np.inf)), np.nextafter(want, np.float32(0))):
            got = corrected(x, s)
            bad = np.flatnonzero(got != want)
            assert bad.size == 0, (n[bad[:5]], got[bad[:5]], want[bad[:5]])
