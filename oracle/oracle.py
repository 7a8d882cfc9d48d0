"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle is a C restatement of EMAN2's "fourier" Reconstructor path
(libEM/reconstructor.cpp, libEM/reconstructor_tools.cpp, libEM/emdata_core.h,
libEM/util.cpp, libEM/transform.cpp, libEM/symmetry.cpp, libEM/ctf.cpp,
libEM/emdata_transform.cpp, libEM/processor.cpp).  Only tests/, the smoke
check and bench.py's cpu_baseline leg import this module; the shipped
library never does.

Array conventions follow the reference:
  slice  : float32 (P, P+2)  rows in wrap order, interleaved re/im
  data   : float32 (P, P, P+2)  [z][y][x re/im]
  norm   : float32 (P, P, P//2+1)
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

GAUSS2 = 0
TRILINEAR = 1
MODES = {"gauss_2": GAUSS2, "trilinear": TRILINEAR, "gauss_var": 2, "gauss_5": 3, "nearest_neighbor": 4, "gauss_3": 5,
         "kaiser_bessel": 6, "gridding_5": 7, "gridding_7": 8,
         "kaiser_bessel_derived": 9,
         "fourier_iter": 10}      # the fourier_iter reconstructor's own insertion (not an inserter mode)

_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)


class CtfStruct(ctypes.Structure):
    _fields_ = [("defocus", ctypes.c_float), ("bfactor", ctypes.c_float),
                ("voltage", ctypes.c_float), ("cs", ctypes.c_float),
                ("ampcont", ctypes.c_float), ("apix", ctypes.c_float)]


def build():
    """Compile both oracle flavours (deterministic and -O3 -ffast-math)."""
    subprocess.check_call(["make", "-s", "-C", _HERE], stdout=subprocess.DEVNULL)


def lib(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name in _LIBS:
        return _LIBS[name]
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    i, f, d, v = ctypes.c_int, ctypes.c_float, ctypes.c_double, None
    sig = {
        "or_fast_exp_table": (v, [_fp]),
        "or_fast_exp": (f, [f]),
        "or_hypot_fast_int": (i, [i, i]),
        "or_hypot_fast": (f, [i, i]),
        "or_round_f": (i, [f]),
        "or_euler_eman": (v, [d, d, d, _fp]),
        "or_transform_mult": (v, [_fp, _fp, _fp]),
        "or_rotation_only": (v, [_fp]),
        "or_sym_nsym": (i, [ctypes.c_char_p]),
        "or_sym_get": (i, [ctypes.c_char_p, _fp]),
        "or_xform2d": (i, [_fp, i, i, _fp, _fp]),
        "or_projection": (v, [i, _fp, _fp, _fp, i, _fp]),
        "or_bessel_i0_scaled": (d, [d]),
        "or_prep_particle_xform": (v, [_fp, i, i, ctypes.POINTER(CtfStruct), _fp, _fp]),
        "or_prep_particle_dfdf": (v, [_fp, i, i, ctypes.POINTER(CtfStruct), f, _fp, _fp]),
        "or_iter_insert": (v, [i, _fp, _fp, _fp, _fp, _fp, i, f, _fp]),
        "or_gauss_fft_project": (i, [i, _fp, _fp, i, i, _fp]),
        "or_iter_ref_from_map": (v, [i, _fp, _fp]),
        "or_insert_slice_ssnr": (i, [i, _fp, _fp, _fp, _fp, _fp, i, _fp, i, i]),
        "or_edge_mean": (f, [_fp, i, i]),
        "or_clip_2d": (v, [_fp, i, i, i, i, i, i, _fp]),
        "or_fft_r2c_1d": (v, [_fp, i, _fp]),
        "or_fft_r2c_2d": (v, [_fp, i, i, _fp]),
        "or_fft_c2r_2d": (v, [_fp, i, i, _fp]),
        "or_fft_r2c_3d": (v, [_fp, i, i, i, _fp]),
        "or_fft_c2r_3d": (v, [_fp, i, i, i, _fp]),
        "or_phaseorigin_tocorner": (v, [_fp, i, i, i]),
        "or_phaseorigin_tocenter": (v, [_fp, i, i, i]),
        "or_ctf_constants": (v, [ctypes.POINTER(CtfStruct), i, _fp, _fp, _fp, _fp]),
        "or_ctf_sign_2d": (v, [ctypes.POINTER(CtfStruct), i, _fp]),
        "or_ctf_amp_2d": (v, [ctypes.POINTER(CtfStruct), i, _fp]),
        "or_preprocess_slice": (v, [_fp, i, _fp]),
        "or_prep_particle": (v, [_fp, i, i, ctypes.POINTER(CtfStruct), _fp]),
        "or_insert_slice_work": (v, [i, _fp, _fp, _fp, _fp, _fp, i, f, i, i]),
        "or_insert_slice_work_f64": (v, [i, _dp, _dp, _fp, _fp, _fp, i, f, i, i]),
        "or_insert_batch_mt": (d, [i, _fp, _fp, _fp, _fp, _fp, i, _fp, i, i, i]),
        "or_normalize_threed": (v, [i, _fp, _fp, i]),
        "or_finish": (v, [i, _fp, _fp, i, i, _fp]),
        "or_determine_slice_agreement": (v, [i, _fp, _fp, _fp, i, _fp, _fp, i, f, i, f, _fp]),
        "or_setup_seed": (v, [i, _fp, _fp, _fp, f]),
    }
    for k, (r, a) in sig.items():
        fn = getattr(L, k)
        fn.restype = r
        fn.argtypes = a
    _LIBS[name] = L
    return L


def _f(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_fp)


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ---------------------------------------------------------------- Util
def fast_exp_table():
    out = np.zeros(1000, np.float32)
    lib().or_fast_exp_table(_f(out))
    return out


def fast_exp(v):
    return lib().or_fast_exp(float(v))


def hypot_fast_int(x, y):
    return lib().or_hypot_fast_int(int(x), int(y))


def hypot_fast(x, y):
    return lib().or_hypot_fast(int(x), int(y))


# ---------------------------------------------------------------- Transform
def euler_eman(az, alt, phi):
    m = np.zeros(12, np.float32)
    lib().or_euler_eman(float(az), float(alt), float(phi), _f(m))
    return m.reshape(3, 4)


def transform_mult(a, b):
    a, b = f32(a).reshape(12), f32(b).reshape(12)
    out = np.zeros(12, np.float32)
    lib().or_transform_mult(_f(a), _f(b), _f(out))
    return out.reshape(3, 4)


def rotation_only(m):
    m = f32(m).reshape(12).copy()
    lib().or_rotation_only(_f(m))
    return m.reshape(3, 4)


def sym_get(sym):
    n = lib().or_sym_nsym(sym.encode())
    if n <= 0:
        raise ValueError("unsupported symmetry %r" % sym)
    out = np.zeros(n * 12, np.float32)
    lib().or_sym_get(sym.encode(), _f(out))
    return out.reshape(n, 3, 4)


# ---------------------------------------------------------------- prep / FFT
def edge_mean(img):
    img = f32(img)
    ny, nx = img.shape
    return lib().or_edge_mean(_f(img), nx, ny)


def fft_r2c_1d(v):
    v = f32(v)
    n = v.shape[0]
    out = np.zeros(2 * (n // 2 + 1), np.float32)
    lib().or_fft_r2c_1d(_f(v), n, _f(out))
    return out


def fft_r2c_2d(img):
    img = f32(img)
    ny, nx = img.shape
    out = np.zeros((ny, 2 * (nx // 2 + 1)), np.float32)
    lib().or_fft_r2c_2d(_f(img), nx, ny, _f(out))
    return out


def fft_c2r_2d(F, nx):
    F = f32(F)
    ny = F.shape[0]
    out = np.zeros((ny, nx), np.float32)
    lib().or_fft_c2r_2d(_f(F), nx, ny, _f(out))
    return out


def fft_r2c_3d(vol):
    vol = f32(vol)
    nz, ny, nx = vol.shape
    out = np.zeros((nz, ny, 2 * (nx // 2 + 1)), np.float32)
    lib().or_fft_r2c_3d(_f(vol), nx, ny, nz, _f(out))
    return out


def fft_c2r_3d(F, nx):
    F = f32(F)
    nz, ny = F.shape[:2]
    out = np.zeros((nz, ny, nx), np.float32)
    lib().or_fft_c2r_3d(_f(F), nx, ny, nz, _f(out))
    return out


def phaseorigin_tocorner(img):
    img = f32(img).copy()
    if img.ndim == 3:
        nz, ny, nx = img.shape
    elif img.ndim == 2:
        nz, (ny, nx) = 1, img.shape
    else:
        nz, ny, nx = 1, 1, img.shape[0]
    lib().or_phaseorigin_tocorner(_f(img), nx, ny, nz)
    return img


def phaseorigin_tocenter(img):
    img = f32(img).copy()
    if img.ndim == 3:
        nz, ny, nx = img.shape
    elif img.ndim == 2:
        nz, (ny, nx) = 1, img.shape
    else:
        nz, ny, nx = 1, 1, img.shape[0]
    lib().or_phaseorigin_tocenter(_f(img), nx, ny, nz)
    return img


def ctf_struct(ctf):
    return CtfStruct(float(ctf["defocus"]), float(ctf.get("bfactor", 0.0)),
                     float(ctf["voltage"]), float(ctf["cs"]),
                     float(ctf["ampcont"]), float(ctf["apix"]))


def ctf_constants(ctf, ny):
    c = ctf_struct(ctf)
    vals = [ctypes.c_float() for _ in range(4)]
    lib().or_ctf_constants(ctypes.byref(c), ny, *[ctypes.byref(v) for v in vals])
    return tuple(v.value for v in vals)  # g1, g2, acac, ds


def ctf_sign_2d(ctf, P):
    out = np.zeros((P, P + 2), np.float32)
    c = ctf_struct(ctf)
    lib().or_ctf_sign_2d(ctypes.byref(c), P, _f(out))
    return out


def ctf_amp_2d(ctf, P):
    out = np.zeros((P, P + 2), np.float32)
    c = ctf_struct(ctf)
    lib().or_ctf_amp_2d(ctypes.byref(c), P, _f(out))
    return out


def preprocess_slice(img):
    img = f32(img)
    P = img.shape[0]
    out = np.zeros((P, P + 2), np.float32)
    lib().or_preprocess_slice(_f(img), P, _f(out))
    return out


def prep_particle(raw, P, ctf=None, xform=None, dfdf=None):
    """caller prep + preprocess_slice; xform: the particle's 3x4 Transform,
    whose 2-D part (translation / scale / mirror) preprocess_slice applies;
    dfdf (with ctf): the flip is the product of CTF_SIGN at the defocus and
    at defocus + dfdf"""
    raw = f32(raw)
    B = raw.shape[0]
    out = np.zeros((P, P + 2), np.float32)
    cp = ctypes.byref(ctf_struct(ctf)) if ctf is not None else None
    if dfdf is not None:
        assert ctf is not None
        xp = None if xform is None else _f(f32(xform).reshape(12).copy())
        lib().or_prep_particle_dfdf(_f(raw), B, P, cp, float(dfdf), xp, _f(out))
        return out
    if xform is None:
        lib().or_prep_particle(_f(raw), B, P, cp, _f(out))
    else:
        lib().or_prep_particle_xform(_f(raw), B, P, cp, _f(f32(xform).reshape(12)), _f(out))
    return out


def gauss_fft_project(vol, xform, mode=0, returnfft=1):
    """GaussFFTProjector::project3d of a real P^3 map: the (P, P+2) Fourier
    slice (returnfft) or the (P, P) image"""
    vol = np.ascontiguousarray(f32(vol))
    P = vol.shape[0]
    out = np.zeros((P, P + 2) if returnfft else (P, P), np.float32)
    rc = lib().or_gauss_fft_project(P, _f(vol), _f(f32(xform).reshape(12).copy()), int(mode), int(returnfft), _f(out))
    assert rc == 0
    return out


def xform2d(img, xform):
    """preprocess_slice's 2-D resampling of a real image; (image, resampled?)"""
    img = f32(img)
    out = np.zeros_like(img)
    r = lib().or_xform2d(_f(img), img.shape[1], img.shape[0], _f(f32(xform).reshape(12)), _f(out))
    return out, bool(r)


# ---------------------------------------------------------------- volume
class Volume:
    """Reference-layout reconstruction state (image + tmp_data)."""

    def __init__(self, P, sym="c1", mode="gauss_2", fp64=False):
        assert P % 2 == 0
        self.P = P
        self.sym = sym
        self.syms = f32(sym_get(sym))
        self.mode = MODES[mode]
        self.fp64 = fp64
        dt = np.float64 if fp64 else np.float32
        self.data = np.zeros((P, P, P + 2), dt)
        self.norm = np.zeros((P, P, P // 2 + 1), dt)

    def insert(self, slice_, rot, weight=1.0, corners=False):
        s = f32(slice_)
        r = f32(rot).reshape(12)
        if self.mode == MODES["fourier_iter"]:
            assert not self.fp64
            ref = getattr(self, "ref", None)
            lib().or_iter_insert(self.P, _f(self.data), _f(self.norm), _f(s), _f(r), _f(self.syms), len(self.syms),
                                 float(weight), None if ref is None else _f(ref))
            return
        if self.fp64:
            lib().or_insert_slice_work_f64(self.P, _d(self.data), _d(self.norm), _f(s), _f(r),
                                          _f(self.syms), len(self.syms), float(weight),
                                          self.mode, int(corners))
        else:
            lib().or_insert_slice_work(self.P, _f(self.data), _f(self.norm), _f(s), _f(r),
                                      _f(self.syms), len(self.syms), float(weight),
                                      self.mode, int(corners))

    def insert_ssnr(self, slice_, rot, ssnr):
        """insert_slice on a usessnr reconstructor of a slice with class_ssnr"""
        assert not self.fp64
        s = f32(slice_)
        r = f32(rot).reshape(12)
        ss = f32(ssnr).ravel()
        rc = lib().or_insert_slice_ssnr(self.P, _f(self.data), _f(self.norm), _f(s), _f(r), _f(self.syms),
                                        len(self.syms), _f(ss), int(ss.size), self.mode)
        assert rc == 0

    def projection(self, euler, ret_fourier=True):
        """FourierReconstructor::projection: (P, P+2) slice or (P, P) image"""
        assert not self.fp64
        P = self.P
        out = np.zeros((P, P + 2) if ret_fourier else (P, P), np.float32)
        lib().or_projection(P, _f(self.data), _f(self.norm), _f(f32(euler).reshape(12)), int(bool(ret_fourier)),
                            _f(out))
        return out

    def setup_iter_ref(self, seed):
        """fourier_iter's setup_seed: the volume cleared, the reference set to
        the complex seed (P, P, P+2) as given, or to do_fft of a real P^3 map
        with the phase origin moved to the corner"""
        P = self.P
        self.data[:] = 0
        self.norm[:] = 0
        seed = f32(seed)
        if seed.shape == (P, P, P):
            ref = np.zeros((P, P, P + 2), np.float32)
            lib().or_iter_ref_from_map(P, _f(np.ascontiguousarray(seed)), _f(ref))
            self.ref = ref
        else:
            assert seed.shape == (P, P, P + 2)
            self.ref = np.ascontiguousarray(seed).copy()

    def setup_seed(self, seed, seed_weight):
        assert not self.fp64
        lib().or_setup_seed(self.P, _f(self.data), _f(self.norm), _f(f32(seed)), float(seed_weight))

    def agreement(self, slice_, rot, weight=1.0, sub=True, apix=1.0):
        assert not self.fp64
        s = f32(slice_).copy()
        r = f32(rot).reshape(12)
        out = np.zeros(5, np.float32)
        lib().or_determine_slice_agreement(self.P, _f(self.data), _f(self.norm), _f(s), self.mode, _f(r),
                                          _f(self.syms), len(self.syms), float(weight),
                                          int(sub), float(apix), _f(out))
        return dict(reconstruct_norm=float(out[0]), reconstruct_absqual=float(out[1]),
                    reconstruct_absqual_lowres=float(out[2]), reconstruct_qual=float(out[3]),
                    reconstruct_weight=float(out[4]))

    def finish(self, doift=True, sqrtnorm=False):
        """Returns the real P^3 map (doift) or the normalised complex volume."""
        data = f32(self.data).copy()
        norm = f32(self.norm).copy()
        out = np.zeros((self.P,) * 3, np.float32) if doift else np.zeros(1, np.float32)
        lib().or_finish(self.P, _f(data), _f(norm), int(sqrtnorm), int(doift), _f(out))
        self.norm_fixed = norm
        return out if doift else data


def insert_batch_mt(P, slices, rots, weights, sym="c1", mode="gauss_2", nthreads=1, fast=True):
    """Multi-threaded batch insertion with private volumes (CPU baseline)."""
    slices = f32(slices)
    rots = f32(rots).reshape(-1, 12)
    weights = f32(weights)
    syms = f32(sym_get(sym))
    data = np.zeros((P, P, P + 2), np.float32)
    norm = np.zeros((P, P, P // 2 + 1), np.float32)
    secs = lib(fast).or_insert_batch_mt(P, _f(data), _f(norm), _f(slices), _f(rots), _f(weights),
                                        len(weights), _f(syms), len(syms), MODES[mode], int(nthreads))
    return data, norm, secs
