"""ctypes binding of libfourier_hip.so (the C ABI in include/fourier_hip.h).

The library is built in-tree by ``python -m eman2_amd.build``.  Loading it
fails loudly when it is missing: there is no CPU fallback for the product
path.  ctypes releases the GIL around each foreign call, as the reference's
Boost.Python wrappers do for insert_slice / finish /
determine_slice_agreement (libpyEM/libpyReconstructor2.cpp:50-83).
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfourier_hip.so")

FH_OK = 0
FH_SKIPPED = -1
FH_HOST = 0
FH_DEVICE = 1

_c_int, _c_float, _c_double, _c_ll = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_longlong
_fp = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p


class Params(ctypes.Structure):
    _fields_ = [("nx", _c_int), ("ny", _c_int), ("nz", _c_int), ("sym", ctypes.c_char_p),
                ("mode", ctypes.c_char_p), ("corners", _c_int), ("sqrtnorm", _c_int),
                ("usessnr", _c_int), ("verbose", _c_int), ("device", _c_int)]


class Ctf(ctypes.Structure):
    _fields_ = [("voltage", _c_float), ("cs", _c_float), ("ampcont", _c_float), ("apix", _c_float),
                ("dfdf", ctypes.c_void_p)]


class Tuning(ctypes.Structure):
    _fields_ = [("waves_per_block", _c_int), ("slices_per_item", _c_int), ("chunk", _c_int), ("blocks_per_cu", _c_int),
                ("graph", _c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("prep_ms", _c_double), ("bin_ms", _c_double), ("insert_ms", _c_double),
                ("total_ms", _c_double), ("insert_kernel_ms", _c_double), ("merge_ms", _c_double), ("slices", _c_ll),
                ("pairs", _c_ll), ("items", _c_ll), ("chunks", _c_ll)]


class Job(ctypes.Structure):
    _fields_ = [("params", Params), ("imgs", _vp), ("n", _c_int), ("box", _c_int), ("where", _c_int),
                ("xforms", _vp), ("weights", _vp), ("defocus", _vp), ("ctf", Ctf), ("use_ctf", _c_int),
                ("out_map", _vp)]


class JobStatus(ctypes.Structure):
    _fields_ = [("state", _c_int), ("error", _c_int), ("t_submit", _c_double), ("t_admit", _c_double),
                ("t_done", _c_double), ("particles_done", _c_ll), ("footprint_bytes", _c_ll)]


JOB_STATES = {0: "queued", 1: "running", 2: "done", 3: "failed", 4: "rejected", 5: "cancelled"}

EXPORTS = {
    "fh_server_create": (_c_int, [_c_int, _c_ll, _c_int, _c_int, ctypes.POINTER(_vp)]),
    "fh_server_submit": (_c_int, [_vp, ctypes.POINTER(Job), ctypes.POINTER(_c_int)]),
    "fh_server_poll": (_c_int, [_vp, _c_int, ctypes.POINTER(JobStatus)]),
    "fh_server_cancel": (_c_int, [_vp, _c_int]),
    "fh_server_drain": (_c_int, [_vp]),
    "fh_server_clock": (_c_double, [_vp]),
    "fh_server_destroy": (_c_int, [_vp]),
    "fh_server_last_error": (ctypes.c_char_p, [_vp]),
    "fh_create": (_c_int, [ctypes.POINTER(Params), ctypes.POINTER(_vp)]),
    "fh_destroy": (_c_int, [_vp]),
    "fh_last_error": (ctypes.c_char_p, [_vp]),
    "fh_setup": (_c_int, [_vp]),
    "fh_setup_seed": (_c_int, [_vp, _vp, _c_float]),
    "fh_setup_seedandweights": (_c_int, [_vp, _vp, _vp]),
    "fh_clear": (_c_int, [_vp]),
    "fh_preprocess_slice": (_c_int, [_vp, _vp, _vp]),
    "fh_insert_slice": (_c_int, [_vp, _vp, _c_int, _vp, _c_float]),
    "fh_insert_batch": (_c_int, [_vp, _vp, _c_int, _c_int, _vp, _vp, _vp, ctypes.POINTER(Ctf), _c_int]),
    "fh_insert_slice_ssnr": (_c_int, [_vp, _vp, _c_int, _vp, _c_float, _vp, _c_int]),
    "fh_projection": (_c_int, [_vp, _vp, _c_int, _vp]),
    "fh_preprocess_slice_xform": (_c_int, [_vp, _vp, _vp, _vp]),
    "fh_insert_batch_ssnr": (_c_int, [_vp, _vp, _c_int, _c_int, _vp, _vp, _vp, _c_int, _vp,
                                      ctypes.POINTER(Ctf), _c_int]),
    "fh_prep_batch": (_c_int, [_vp, _vp, _c_int, _c_int, _vp, ctypes.POINTER(Ctf), _vp, _c_int]),
    "fh_insert_plain": (_c_int, [_vp, _vp, _c_int, _c_int, _vp, _vp, _vp, ctypes.POINTER(Ctf)]),
    "fh_insert_plain_where": (_c_int, [_vp, _vp, _c_int, _c_int, _vp, _vp, _vp, ctypes.POINTER(Ctf), _c_int]),
    "fh_determine_slice_agreement": (_c_int, [_vp, _vp, _c_int, _vp, _c_float, _c_int, _c_float, _vp]),
    "fh_determine_slice_agreement_ssnr": (_c_int, [_vp, _vp, _c_int, _vp, _c_float, _c_int, _c_float, _vp, _c_int,
                                                   _vp]),
    "fh_setup_seed_map": (_c_int, [_vp, _vp]),
    "fh_gauss_fft_volume": (_c_int, [_vp, _vp]),
    "fh_gauss_fft_project": (_c_int, [_vp, _c_int, _vp, _c_int, _c_int, _vp]),
    "fh_merge_add": (_c_int, [_vp, _vp, _vp, _c_int]),
    "fh_merge_finish": (_c_int, [_vp, _vp]),
    "fh_determine_slice_agreement_noalloc": (_c_int, [_vp, _vp, _vp, _c_float, _c_int, _c_float, _vp]),
    "fh_agreement_batch": (_c_int, [_vp, _vp, _c_int, _c_int, _vp, _vp, _vp, ctypes.POINTER(Ctf), _c_int,
                                    _c_float, _vp, _c_int]),
    "fh_finish": (_c_int, [_vp, _c_int, _vp, _vp]),
    "fh_get_state": (_c_int, [_vp, _vp, _vp]),
    "fh_set_tuning": (_c_int, [_vp, ctypes.POINTER(Tuning)]),
    "fh_get_tuning": (_c_int, [_vp, ctypes.POINTER(Tuning)]),
    "fh_get_stats": (_c_int, [_vp, ctypes.POINTER(Stats)]),
    "fh_sync": (_c_int, [_vp]),
    "fh_footprint_bytes": (_c_ll, [_c_int, _c_int, _c_int]),
    "fh_get_symmetries": (_c_int, [ctypes.c_char_p, ctypes.c_void_p, _c_int]),
}

_LIB = None


def load(path=None):
    """Load and type the library (raises if it is not built)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError("libfourier_hip.so is not built (expected at %s); run python -m eman2_amd.build" % p)
    L = ctypes.CDLL(p)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = L
    return L


# status code -> exception class name of the reference (libEM/exception.h)
ERRORS = {
    1: "InvalidParameterException",
    2: "NotExistingObjectException",
    3: "ImageDimensionException",
    4: "NullPointerException",
    5: "DeviceError",
    6: "StateError",
    7: "OutOfMemoryError",
}


class FourierHipError(RuntimeError):
    """Raised for a non-zero status; the reference raises its E2Exception
    family, which surfaces in Python as RuntimeError."""

    def __init__(self, code, msg):
        self.code = code
        super().__init__("%s: %s" % (ERRORS.get(code, "Error%d" % code), msg))


def check(rc, handle=None):
    if rc == FH_OK or rc == FH_SKIPPED:
        return rc
    msg = load().fh_last_error(handle)
    raise FourierHipError(rc, msg.decode() if msg else "")
