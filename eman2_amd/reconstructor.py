"""Python surface of the fourier_hip reconstructor, mirroring libpyEM's
``Reconstructors`` / ``Reconstructor`` binding
(libpyEM/libpyReconstructor2.cpp:192-224) so e2make3dpar-style callers
switch by name:

    recon = Reconstructors.get("fourier_hip", {"size": (P, P, P), "sym": "c1",
                                               "mode": "gauss_2"})
    recon.setup()
    recon.insert_slice(img, Transform({"type": "eman", "az": .., ...}), 1.0)
    vol = recon.finish(True)

Parameter validation follows Factory<T>::get (libEM/emobject.h:801-834): an
unknown name raises NotExistingObjectException, an unknown key
InvalidParameterException (both RuntimeError subclasses here, as in the
reference's Python binding).  Every call goes to the HIP library; there is
no CPU path.
"""
import ctypes

import numpy as np

from . import _lib
from .transform import Transform

TUNING_KEYS = ("waves_per_block", "slices_per_item", "chunk", "blocks_per_cu", "graph")


class EMData:
    """Numpy-backed stand-in for the EMData objects the reconstructor API
    exchanges: a real image/volume, or a complex Fourier slice/volume in the
    reference's interleaved layout (rows of nx+2 floats), plus attributes."""

    def __init__(self, array=None, complex_=False, attrs=None):
        self._a = None if array is None else np.ascontiguousarray(array, np.float32)
        self._complex = complex_
        self.attr = dict(attrs or {})

    @classmethod
    def from_numpy(cls, a):
        return cls(a)

    def numpy(self):
        return self._a

    def is_complex(self):
        return self._complex

    def get_xsize(self):
        return self._a.shape[-1]

    def get_ysize(self):
        return self._a.shape[-2] if self._a.ndim > 1 else 1

    def get_zsize(self):
        return self._a.shape[0] if self._a.ndim > 2 else 1

    def get_attr(self, k):
        return self.attr[k]

    def get_attr_default(self, k, d):
        return self.attr.get(k, d)

    def set_attr(self, k, v):
        self.attr[k] = v

    def has_attr(self, k):
        return k in self.attr

    def __getitem__(self, k):
        return self.attr[k]

    def __setitem__(self, k, v):
        self.attr[k] = v

    def mult(self, f):
        self._a = (self._a * np.float32(f)).astype(np.float32)


def _arr(x):
    if isinstance(x, EMData):
        return x.numpy()
    return np.ascontiguousarray(x, np.float32)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _xform12(t):
    if isinstance(t, Transform):
        m = t.matrix
    else:
        m = np.asarray(t, np.float32)
        if m.shape == (3, 3):
            m = np.concatenate([m, np.zeros((3, 1), np.float32)], 1)
    return np.ascontiguousarray(m, np.float32).reshape(12)


class FourierHipReconstructor:
    """The "fourier_hip" Reconstructor (FourierReconstructor semantics,
    libEM/reconstructor.cpp:686-1641) on an MI355X."""

    NAME = "fourier_hip"

    # the "fourier" TypeDict (libEM/reconstructor.h:501-517) plus this
    # build's own keys (pad, device, tuning)
    PARAM_TYPES = {
        "size": "INTARRAY", "sym": "STRING", "mode": "STRING", "corners": "BOOL",
        "sqrtnorm": "BOOL", "usessnr": "BOOL", "verbose": "BOOL", "quiet": "BOOL",
        "subvolume": "INTARRAY", "savenorm": "STRING", "normout": "EMDATA",
        "pad": "INT", "device": "INT", "waves_per_block": "INT", "slices_per_item": "INT",
        "chunk": "INT", "blocks_per_cu": "INT",
    }

    def __init__(self, params=None):
        self.params = {}
        self._h = None
        self._lib = _lib.load()
        self.set_params(params or {})

    # ---------------------------------------------------------------- params
    def get_name(self):
        return self.NAME

    def get_desc(self):
        return "Direct Fourier inversion (gauss_2/trilinear) on MI355X"

    def get_param_types(self):
        return dict(self.PARAM_TYPES)

    def get_params(self):
        return dict(self.params)

    def set_params(self, params):
        for k in params:
            if k not in self.PARAM_TYPES:
                raise _lib.FourierHipError(1, "%s is not a valid parameter for %s" % (k, self.NAME))
        if "subvolume" in params:
            raise _lib.FourierHipError(1, "subvolume is not supported by %s" % self.NAME)
        construct = ("size", "sym", "mode", "corners", "sqrtnorm", "usessnr", "verbose", "device")
        changed = self._h is None or any(k in construct and params[k] != self.params.get(k) for k in params)
        self.params.update(params)
        if changed:
            self._create()
        elif any(k in params for k in TUNING_KEYS):
            t = self.get_tuning()
            t.update({k: int(params[k]) for k in TUNING_KEYS if k in params})
            self.set_tuning(**t)

    def set_param(self, k, v):
        self.set_params({k: v})

    def insert_params(self, p):
        self.set_params(p)

    def _create(self):
        p = self.params
        if "size" not in p:
            return
        size = p["size"]
        if np.isscalar(size):
            size = (size, size, size)
        nx, ny, nz = (int(s) for s in size)
        if self._h is not None:
            self._lib.fh_destroy(self._h)
            self._h = None
        prm = _lib.Params(nx, ny, nz, str(p.get("sym", "c1")).encode(), str(p.get("mode", "gauss_2")).encode(),
                          int(bool(p.get("corners", False))), int(bool(p.get("sqrtnorm", False))),
                          int(bool(p.get("usessnr", False))), int(bool(p.get("verbose", False))),
                          int(p.get("device", 0)))
        h = ctypes.c_void_p()
        rc = self._lib.fh_create(ctypes.byref(prm), ctypes.byref(h))
        _lib.check(rc, None)
        self._h = h
        self.P = nx
        from .autotune import lookup, nsym_of
        try:
            best = lookup(nx, nsym_of(str(p.get("sym", "c1"))))
        except (ValueError, IndexError):
            best = None
        if best:
            self.set_tuning(**best)
        if any(k in p for k in TUNING_KEYS):
            t = self.get_tuning()
            t.update({k: int(p[k]) for k in TUNING_KEYS if k in p})
            self.set_tuning(**t)

    def __del__(self):
        try:
            if self._h is not None:
                self._lib.fh_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def _need(self):
        if self._h is None:
            raise _lib.FourierHipError(1, "size is a required parameter")
        return self._h

    def _check(self, rc):
        return _lib.check(rc, self._h)

    # ---------------------------------------------------------------- lifecycle
    def setup(self):
        self._check(self._lib.fh_setup(self._need()))

    def setup_seed(self, seed, seed_weight):
        s = _arr(seed)
        self._check(self._lib.fh_setup_seed(self._need(), _ptr(s), float(seed_weight)))

    def setup_seedandweights(self, seed, weights):
        s, w = _arr(seed), _arr(weights)
        self._check(self._lib.fh_setup_seedandweights(self._need(), _ptr(s), _ptr(w)))

    def clear(self):
        self._check(self._lib.fh_clear(self._need()))

    # ---------------------------------------------------------------- slices
    def preprocess_slice(self, img, xform=None):
        """Reconstructor::preprocess_slice: returns the slice (EMData, complex
        layout (P, P+2)) with attribute reconstruct_preproc=1."""
        a = _arr(img)
        P = self.P
        if a.shape != (P, P):
            raise _lib.FourierHipError(3, "image must be %dx%d" % (P, P))
        out = np.zeros((P, P + 2), np.float32)
        if xform is None:
            self._check(self._lib.fh_preprocess_slice(self._need(), _ptr(a), _ptr(out)))
        else:  # the Transform's 2-D part (translation / scale / mirror) resamples the image first
            m = _xform12(xform)
            self._check(self._lib.fh_preprocess_slice_xform(self._need(), _ptr(a), _ptr(m), _ptr(out)))
        e = EMData(out, complex_=True, attrs={"reconstruct_preproc": 1})
        if isinstance(img, EMData):
            for k, v in img.attr.items():
                e.attr.setdefault(k, v)
        return e

    def insert_slice(self, img, xform, weight=1.0):
        pre = int(isinstance(img, EMData) and bool(img.get_attr_default("reconstruct_preproc", 0)))
        a = _arr(img)
        P = self.P
        want = (P, P + 2) if pre else (P, P)
        if a.shape != want:
            raise _lib.FourierHipError(3, "slice must be %s, got %s" % (want, a.shape))
        m = _xform12(xform)
        if isinstance(img, EMData) and img.has_attr("class_ssnr"):
            # usessnr weights the slice by its class_ssnr (reconstructor.cpp:1080-1088)
            ss = np.ascontiguousarray(img.get_attr("class_ssnr"), np.float32).ravel()
            return self._check(self._lib.fh_insert_slice_ssnr(self._need(), _ptr(a), pre, _ptr(m), float(weight),
                                                              _ptr(ss), int(ss.size)))
        return self._check(self._lib.fh_insert_slice(self._need(), _ptr(a), pre, _ptr(m), float(weight)))

    def _batch_args(self, imgs, xforms, weights, defocus, ctf, nprep=None):
        n = len(xforms)
        nd = n if nprep is None else nprep
        if isinstance(xforms, np.ndarray) and xforms.ndim == 3 and xforms.shape[1:] == (3, 4):
            m = np.ascontiguousarray(xforms, np.float32).reshape(n, 12)
        elif isinstance(xforms, np.ndarray) and xforms.ndim == 2 and xforms.shape[1] == 12:
            m = np.ascontiguousarray(xforms, np.float32)
        else:
            m = np.ascontiguousarray(np.stack([_xform12(t) for t in xforms]) if n else np.zeros((0, 12)), np.float32)
        w = np.ascontiguousarray(np.ones(n) if weights is None else weights, np.float32)
        d = None if defocus is None else np.ascontiguousarray(defocus, np.float32)
        c = None
        if ctf is not None:
            c = _lib.Ctf(float(ctf.get("voltage", 300.0)), float(ctf.get("cs", 2.7)),
                         float(ctf.get("ampcont", 10.0)), float(ctf["apix"]))
            if ctf.get("dfdf") is not None:
                # per-particle defocus offsets: the flip is sign(defocus) * sign(defocus + dfdf)
                dd = np.ascontiguousarray(ctf["dfdf"], np.float32).reshape(-1)
                if dd.shape[0] != nd:
                    raise _lib.FourierHipError(1, "ctf['dfdf'] must hold one offset per particle")
                c._keep = dd
                c.dfdf = dd.ctypes.data
        return n, m, w, d, c

    def insert_batch(self, imgs, xforms, weights=None, defocus=None, ctf=None, box=None, class_ssnr=None):
        """Batched insertion of raw particles (numpy (n,box,box) on the host,
        or a device pointer (int) with ``box`` given) with the caller prep
        fused in (edge-mean subtract, pad to size, optional CTF phase flip).
        ``class_ssnr`` (n, len): every particle's class_ssnr attribute, used
        for per-radius weights when the reconstructor has ``usessnr``."""
        n, m, w, d, c = self._batch_args(imgs, xforms, weights, defocus, ctf)
        ss = None
        if class_ssnr is not None:
            ss = np.ascontiguousarray(class_ssnr, np.float32)
            if ss.ndim != 2 or ss.shape[0] != n:
                raise _lib.FourierHipError(1, "class_ssnr must be (n, len) for n particles")
        if isinstance(imgs, int):
            where, ptr, B = _lib.FH_DEVICE, ctypes.c_void_p(imgs), int(box)
        else:
            a = _arr(imgs)
            if a.shape[0] != n:
                raise _lib.FourierHipError(3, "image / transform count mismatch")
            where, ptr, B = _lib.FH_HOST, _ptr(a), a.shape[-1]
        if ss is not None:
            return self._check(self._lib.fh_insert_batch_ssnr(self._need(), ptr, n, B, _ptr(m), _ptr(w), _ptr(ss),
                                                              int(ss.shape[1]), None if d is None else _ptr(d),
                                                              None if c is None else ctypes.byref(c), where))
        return self._check(self._lib.fh_insert_batch(self._need(), ptr, n, B, _ptr(m), _ptr(w),
                                                     None if d is None else _ptr(d),
                                                     None if c is None else ctypes.byref(c), where))

    def prep_batch(self, imgs, defocus=None, ctf=None):
        """Batched caller prep + preprocess_slice: (n,box,box) -> (n,P,P+2)."""
        a = _arr(imgs)
        n = a.shape[0]
        _, _, _, d, c = self._batch_args(None, [None] * 0, None, defocus, ctf, nprep=n)
        out = np.zeros((n, self.P, self.P + 2), np.float32)
        self._check(self._lib.fh_prep_batch(self._need(), _ptr(a), n, a.shape[-1],
                                            None if d is None else _ptr(d),
                                            None if c is None else ctypes.byref(c), _ptr(out), _lib.FH_HOST))
        return out

    def insert_plain(self, imgs, xforms, weights=None, defocus=None, ctf=None, box=None):
        """Same contract as insert_batch, one particle per launch (baseline);
        imgs a numpy stack (copied per particle) or a device pointer with box."""
        n, m, w, d, c = self._batch_args(imgs, xforms, weights, defocus, ctf)
        if isinstance(imgs, int):
            where, ptr, B = _lib.FH_DEVICE, ctypes.c_void_p(imgs), int(box)
        else:
            a = _arr(imgs)
            where, ptr, B = _lib.FH_HOST, _ptr(a), a.shape[-1]
        return self._check(self._lib.fh_insert_plain_where(self._need(), ptr, n, B, _ptr(m), _ptr(w),
                                                           None if d is None else _ptr(d),
                                                           None if c is None else ctypes.byref(c), where))

    def projection(self, xform, ret_fourier=0):
        """FourierReconstructor::projection (libEM/reconstructor.cpp:1383-1445):
        a new EMData, the complex (P, P+2) slice when ret_fourier, else the
        real (P, P) image."""
        m = _xform12(xform)
        P = self.P
        out = np.zeros((P, P + 2) if ret_fourier else (P, P), np.float32)
        self._check(self._lib.fh_projection(self._need(), _ptr(m), int(bool(ret_fourier)), _ptr(out)))
        return EMData(out, bool(ret_fourier))

    def determine_slice_agreement(self, img, xform, weight=1.0, sub=True, reference_sub=False):
        """sub: the slice's own contribution is removed before the compare and
        added back after it.  reference_sub: the reference's own sequence
        instead (reconstructor.cpp:1212-1257) -- its removal call passes
        -weight, the SSNR flag of do_insert_slice_work, so the slice is added
        with its class_ssnr weights, compared, and added again with +weight
        (the volume keeps both); needs the slice's class_ssnr attribute
        (NotExistingObjectException without it, as in the reference)."""
        pre = int(isinstance(img, EMData) and bool(img.get_attr_default("reconstruct_preproc", 0)))
        a = _arr(img)
        m = _xform12(xform)
        out = np.zeros(5, np.float32)
        apix = float(img.get_attr_default("apix_x", 1.0)) if isinstance(img, EMData) else 1.0
        if sub and reference_sub:
            has = isinstance(img, EMData) and img.has_attr("class_ssnr")
            ss = np.ascontiguousarray(img.get_attr("class_ssnr"), np.float32).ravel() if has else None
            self._check(self._lib.fh_determine_slice_agreement_ssnr(
                self._need(), _ptr(a), pre, _ptr(m), float(weight), 1, apix, None if ss is None else _ptr(ss),
                0 if ss is None else int(ss.size), _ptr(out)))
        else:
            self._check(self._lib.fh_determine_slice_agreement(self._need(), _ptr(a), pre, _ptr(m), float(weight),
                                                               int(bool(sub)), apix, _ptr(out)))
        keys = ("reconstruct_norm", "reconstruct_absqual", "reconstruct_absqual_lowres", "reconstruct_qual",
                "reconstruct_weight")
        res = dict(zip(keys, (float(v) for v in out)))
        if isinstance(img, EMData):
            for k, v in res.items():
                if k != "reconstruct_qual":
                    img.set_attr(k, v)
        return res

    def agreement_batch(self, imgs, xforms, weights=None, defocus=None, ctf=None, sub=True, apix=1.0, box=None,
                        out=None):
        """determine_slice_agreement for a batch of raw particles: numpy
        (n,box,box) on the host, returning (n, 5) = norm, absqual,
        absqual_lowres, qual, weight; or a device pointer (int) with ``box``
        and ``out`` a device pointer to n x 5 floats, filled in place."""
        n, m, w, d, c = self._batch_args(imgs, xforms, weights, defocus, ctf)
        if isinstance(imgs, int):
            if out is None:
                raise _lib.FourierHipError(3, "device particles need a device output pointer")
            where, ptr, B, optr, res = _lib.FH_DEVICE, ctypes.c_void_p(imgs), int(box), ctypes.c_void_p(out), None
        else:
            a = _arr(imgs)
            res = np.zeros((n, 5), np.float32)
            where, ptr, B, optr = _lib.FH_HOST, _ptr(a), a.shape[-1], _ptr(res)
        self._check(self._lib.fh_agreement_batch(self._need(), ptr, n, B, _ptr(m), _ptr(w),
                                                 None if d is None else _ptr(d),
                                                 None if c is None else ctypes.byref(c), int(bool(sub)),
                                                 float(apix), optr, where))
        return res

    # ---------------------------------------------------------------- finish
    def finish(self, doift=True, return_numpy=False):
        P = self.P
        out = np.zeros((P, P, P) if doift else (P, P, P + 2), np.float32)
        norm = np.zeros((P, P, P // 2 + 1), np.float32)
        self._check(self._lib.fh_finish(self._need(), int(bool(doift)), _ptr(out), _ptr(norm)))
        normout = self.params.get("normout")
        if isinstance(normout, EMData):
            normout._a = norm
        elif isinstance(normout, np.ndarray) and normout.shape == norm.shape:
            normout[...] = norm
        self.last_norm = norm
        if return_numpy:
            return out
        return EMData(out, complex_=not doift)

    # ---------------------------------------------------------------- task merge
    def merge_add(self, threed, normout):
        """e2spa_make3d.py:180-184 on this (set up) reconstructor: a task's
        finish(False) volume times its normout added to the sum, the normout
        to the summed weights"""
        t = _arr(threed.numpy() if isinstance(threed, EMData) else threed)
        w = _arr(normout.numpy() if isinstance(normout, EMData) else normout)
        P = self.P
        if t.shape != (P, P, P + 2) or w.shape != (P, P, P // 2 + 1):
            raise _lib.FourierHipError(3, "merge_add: volume (P, P, P+2) and normout (P, P, P/2+1) expected")
        self._check(self._lib.fh_merge_add(self._need(), _ptr(t), _ptr(w), _lib.FH_HOST))

    def merge_finish(self, return_numpy=False):
        """e2spa_make3d.py:186-192: the merged map (P^3, phase origin centred)"""
        P = self.P
        out = np.zeros((P, P, P), np.float32)
        self._check(self._lib.fh_merge_finish(self._need(), _ptr(out)))
        return out if return_numpy else EMData(out)

    def get_state(self):
        P = self.P
        data = np.zeros((P, P, P + 2), np.float32)
        norm = np.zeros((P, P, P // 2 + 1), np.float32)
        self._check(self._lib.fh_get_state(self._need(), _ptr(data), _ptr(norm)))
        return data, norm

    # ---------------------------------------------------------------- tuning
    def set_tuning(self, waves_per_block=2, slices_per_item=1024, chunk=0, blocks_per_cu=8, graph=0):
        t = _lib.Tuning(int(waves_per_block), int(slices_per_item), int(chunk), int(blocks_per_cu), int(graph))
        self._check(self._lib.fh_set_tuning(self._need(), ctypes.byref(t)))

    def get_tuning(self):
        t = _lib.Tuning()
        self._check(self._lib.fh_get_tuning(self._need(), ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in _lib.Tuning._fields_}

    def get_stats(self):
        s = _lib.Stats()
        self._check(self._lib.fh_get_stats(self._need(), ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in _lib.Stats._fields_}

    def sync(self):
        self._check(self._lib.fh_sync(self._need()))


class FourierIterHipReconstructor(FourierHipReconstructor):
    """The "fourier_iter_hip" Reconstructor: FourierIterReconstructor
    (libEM/reconstructor.cpp:428-682) on an MI355X.  Its own 6x6x6 Gaussian
    insertion (no inserter modes); setup_seed with the previous map makes
    that map the reference volume of the next pass (e2make3dpar.py
    --iterative); determine_slice_agreement returns without comparing."""

    NAME = "fourier_iter_hip"
    PARAM_TYPES = {
        "size": "INTARRAY", "sym": "STRING", "verbose": "BOOL", "quiet": "BOOL", "savenorm": "STRING",
        "normout": "EMDATA", "pad": "INT", "device": "INT", "waves_per_block": "INT", "slices_per_item": "INT",
        "chunk": "INT", "blocks_per_cu": "INT",
    }

    def __init__(self, params=None):
        super().__init__(params)

    def set_params(self, params):
        super().set_params(params)
        self.params["mode"] = "fourier_iter"

    def _create(self):
        self.params["mode"] = "fourier_iter"
        super()._create()

    def get_desc(self):
        return "Iterative direct Fourier inversion (fourier_iter) on MI355X"

    def setup_seed(self, seed, seed_weight=1.0):
        """a real P^3 map (the previous pass's finish(True)) becomes the
        reference after do_fft and the phase origin moved to the corner, on
        the GPU; a complex (P, P, P+2) volume is taken as it is
        (reconstructor.cpp:482-489).  The seed weight is not used."""
        a = _arr(seed)
        P = self.P
        if a.shape == (P, P, P) and not (isinstance(seed, EMData) and seed.is_complex()):
            self._check(self._lib.fh_setup_seed_map(self._need(), _ptr(a)))
        else:
            self._check(self._lib.fh_setup_seed(self._need(), _ptr(a), float(seed_weight)))

    def setup_seedandweights(self, seed, weights):
        self.setup_seed(seed, 1.0)

    def determine_slice_agreement(self, img, xform, weight=1.0, sub=True, reference_sub=False):
        """returns 0 without comparing (reconstructor.cpp:639): no attributes set"""
        self._need()
        return {}


class Reconstructors:
    """Factory<Reconstructor> (libEM/emobject.h:724-889) with
    "fourier_hip" and "fourier_iter_hip" registered; ``add`` mirrors
    Factory<T>::add<>()."""

    _registry = {FourierHipReconstructor.NAME: FourierHipReconstructor,
                 FourierIterHipReconstructor.NAME: FourierIterHipReconstructor}

    @classmethod
    def get(cls, name, params=None):
        if name not in cls._registry:
            raise _lib.FourierHipError(2, "No such an instance existing: %s" % name)
        return cls._registry[name](params or {})

    @classmethod
    def get_list(cls):
        return sorted(cls._registry)

    @classmethod
    def add(cls, klass):
        cls._registry[klass.NAME] = klass
