"""fh_determine_slice_agreement_noalloc (the allocation-free entry point)
against fh_determine_slice_agreement on the same volumes, through the C ABI:
two calls in a row on one GPU, the second on a smaller volume (the shared
scratch keeps the first call's contents), with and without the slice's own
contribution removed, a 2x2x2 and a 6x6x6 inserter; outputs within 1e-4
relative, the volume restored; sizes beyond the scratch refused."""
import ctypes

import numpy as np
import pytest

from eman2_amd import synthetic as S
from eman2_amd.transform import Transform

pytestmark = pytest.mark.gpu
KEYS = ("reconstruct_norm", "reconstruct_absqual", "reconstruct_absqual_lowres", "reconstruct_qual",
        "reconstruct_weight")


def _hip():
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64" in line}
    assert paths, "libamdhip64 not loaded"
    rocm = sorted(q for q in paths if q.startswith("/opt/rocm"))
    hip = ctypes.CDLL((rocm or sorted(paths))[0])
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def _case(oracle, P, sym, mode, n, seed):
    from eman2_amd import Reconstructors
    print("noalloc case P=%d sym=%s mode=%s" % (P, sym, mode), flush=True)
    B = P * 3 // 4
    ph = S.phantom(10, B * 0.3, seed=seed)
    imgs, mats, _ = S.particles(ph, n, B, seed=seed + 1, snr=0.5)
    r = Reconstructors.get("fourier_hip", dict(size=(P, P, P), sym=sym, mode=mode))
    r.setup()
    r.insert_batch(imgs, list(mats), np.ones(n, np.float32))
    slices = [oracle.prep_particle(imgs[i], P) for i in range(3)]
    return r, slices, mats


@pytest.mark.parametrize("first,second", [((64, "c1", "gauss_2", 900), (32, "d2", "gauss_2", 500)),
                                          ((48, "d2", "gauss_var", 300), (32, "c4", "trilinear", 400))])
def test_noalloc_agreement_matches_allocating_entry_point(oracle, first, second):
    from eman2_amd import _lib
    from eman2_amd.reconstructor import EMData
    L = _lib.load()
    hip = _hip()
    for k, (P, sym, mode, n) in enumerate((first, second)):
        r, slices, mats = _case(oracle, P, sym, mode, n, 10 * P + k)
        for i, sub in ((0, True), (1, False), (2, True)):
            sl = np.ascontiguousarray(slices[i], np.float32)
            d = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(d), sl.nbytes) == 0
            try:
                assert hip.hipMemcpy(d, sl.ctypes.data, sl.nbytes, 1) == 0
                before = r.get_state()
                out = np.zeros(5, np.float32)
                m = np.ascontiguousarray(mats[i], np.float32).reshape(12)
                rc = L.fh_determine_slice_agreement_noalloc(r._h, d, m.ctypes.data, ctypes.c_float(0.8), int(sub),
                                                            ctypes.c_float(1.3), out.ctypes.data)
                assert rc == 0, L.fh_last_error(r._h)
                after = r.get_state()
                for a, b in zip(after, before):
                    assert np.abs(a - b).max() <= 1e-5 * max(1.0, np.abs(b).max())
                ref = r.determine_slice_agreement(EMData(sl, True, {"reconstruct_preproc": 1, "apix_x": 1.3}),
                                                  Transform(mats[i]), 0.8, sub)
            finally:
                hip.hipFree(d)
            for j, key in enumerate(KEYS):
                assert abs(out[j] - ref[key]) <= 1e-4 * max(1.0, abs(ref[key])), (P, i, key, out[j], ref[key])


def test_noalloc_refuses_sizes_beyond_scratch():
    """icos at pad 600: 708 compare blocks x 60 operators x 64 B > 2 MiB"""
    from eman2_amd import Reconstructors, _lib
    L = _lib.load()
    r = Reconstructors.get("fourier_hip", dict(size=(600, 600, 600), sym="icos", mode="gauss_2"))
    r.setup()
    out = np.zeros(5, np.float32)
    m = np.eye(3, 4, dtype=np.float32).reshape(12)
    hip = _hip()
    d = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d), 4 * 600 * 602) == 0
    try:
        assert hip.hipMemcpy(d, np.zeros((600, 602), np.float32).ctypes.data, 4 * 600 * 602, 1) == 0
        rc = L.fh_determine_slice_agreement_noalloc(r._h, d, m.ctypes.data, ctypes.c_float(1.0), 1,
                                                    ctypes.c_float(1.0), out.ctypes.data)
    finally:
        hip.hipFree(d)
    assert rc == 1  # FH_ERR_INVALID_PARAM
    assert b"2 MiB" in L.fh_last_error(r._h)
