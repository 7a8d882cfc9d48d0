"""The gauss_fft projector on the GPU, mirroring the reference's Projector
factory (Factory<Projector>, libEM/projector.cpp:57-66) and
EMData.project("gauss_fft", {...}) (GaussFFTProjector::project3d,
libEM/projector.cpp:69-295), as e2make3dpar.py:716-721 calls it:

    pj = Projectors.get("gauss_fft", {"transform": xf, "returnfft": 1}).project3d(ref_map)

Modes 0 (trilinear, the default), 1 (nearest) and 2 (Gaussian).  The map's
transform is computed once per volume on the GPU and kept there, so
``project_batch(volume, xforms)`` makes many projections of one map in one
call.  There is no CPU path.
"""
import ctypes

import numpy as np

from . import _lib
from .reconstructor import EMData, FourierHipReconstructor, _xform12


class GaussFFTProjector:
    NAME = "gauss_fft"
    PARAM_TYPES = {"transform": "TRANSFORM", "mode": "INT", "returnfft": "INT", "device": "INT"}

    def __init__(self, params=None):
        self.params = {}
        self._ctx = None          # a fourier_hip handle of the volume's size: the GPU context
        self._vol_key = None
        self.set_params(params or {})

    def set_params(self, params):
        for k in params:
            if k not in self.PARAM_TYPES:
                raise _lib.FourierHipError(1, "%s is not a valid parameter for %s" % (k, self.NAME))
        self.params.update(params)

    def get_name(self):
        return self.NAME

    def _context(self, vol):
        P = vol.shape[0]
        if vol.shape != (P, P, P) or P % 2:
            raise _lib.FourierHipError(3, "gauss_fft projects an even cubic real volume")
        if self._ctx is None or self._ctx.P != P:
            self._ctx = FourierHipReconstructor({"size": (P, P, P), "device": int(self.params.get("device", 0))})
            self._vol_key = None
        key = (vol.__array_interface__["data"][0], vol.shape, float(vol.ravel()[::max(1, vol.size // 64)].sum()))
        if key != self._vol_key:
            self._ctx._check(self._ctx._lib.fh_gauss_fft_volume(self._ctx._need(),
                                                                 vol.ctypes.data_as(ctypes.c_void_p)))
            self._vol_key = key
        return self._ctx

    def project_batch(self, volume, xforms, mode=None, returnfft=None):
        """n projections of one real P^3 map: (n, P, P+2) Fourier slices or (n, P, P) images"""
        vol = volume.numpy() if isinstance(volume, EMData) else np.ascontiguousarray(volume, np.float32)
        if isinstance(volume, EMData) and volume.is_complex():
            raise _lib.FourierHipError(1, "gauss_fft on a complex volume is not supported by fourier_hip")
        vol = np.ascontiguousarray(vol, np.float32)
        ctx = self._context(vol)
        mode = int(self.params.get("mode", 0) if mode is None else mode)
        rf = int(self.params.get("returnfft", 0) if returnfft is None else returnfft)
        m = np.ascontiguousarray(np.stack([_xform12(t) for t in xforms]) if len(xforms) else np.zeros((0, 12)),
                                 np.float32)
        P = vol.shape[0]
        out = np.zeros((len(m), P, P + 2) if rf else (len(m), P, P), np.float32)
        ctx._check(ctx._lib.fh_gauss_fft_project(ctx._need(), len(m), m.ctypes.data_as(ctypes.c_void_p), mode, rf,
                                                 out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def project3d(self, volume):
        t = self.params.get("transform")
        if t is None:
            raise _lib.FourierHipError(5, "The transform object (required for projection), was not specified")
        rf = int(self.params.get("returnfft", 0))
        out = self.project_batch(volume, [t])[0]
        e = EMData(out, complex_=bool(rf))
        e.set_attr("xform.projection", t)
        return e


class Projectors:
    """Factory<Projector> with "gauss_fft" registered"""

    _registry = {GaussFFTProjector.NAME: GaussFFTProjector}

    @classmethod
    def get(cls, name, params=None):
        if name not in cls._registry:
            raise _lib.FourierHipError(2, "No such an instance existing: %s" % name)
        return cls._registry[name](params or {})

    @classmethod
    def get_list(cls):
        return sorted(cls._registry)
