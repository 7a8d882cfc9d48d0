"""EMAN2's HDF5 image files (libEM/io/hdfio2.cpp) read and written through
the HDF5 C library (ctypes; h5py is not needed).

Layout (hdfio2.cpp:427-480, 1266-1368, 1415-1610): the group /MDF/images
carries the int attribute ``imageid_max`` (the highest image number); image
n is the group /MDF/images/<n> with the image's header as attributes named
"EMAN.<key>" and one dataset ``image`` of shape (nz, ny, nx), (ny, nx) or
(nx,).  Attribute types follow HdfIO2::write_attr (:234-425): bool as the
char 'T'/'F', int, unsigned int, float, double, strings, 1-D float / int
arrays, a Transform as a compound of twelve floats "00" .. "23".

Bit-reduced ("compressed") images (write_compressed, libpyEM/EMAN2.py:
3142-3258; Renderer::getRenderedDataAndRendertrunc, libEM/io/renderer.h:
74-110; EMUtil::getRenderMinMax, libEM/emutil.cpp:1739-1860) store
round((v - rmin) / (rmax - rmin) * (2^bits - 1)) as uint8 (bits <= 8) or
uint16, chunked (1, 8, nx) with deflate, and the group attributes
EMAN.stored_rendermin / stored_rendermax / stored_renderbits /
stored_truncated; reading rescales v = (i / (2^bits - 1)) (rmax - rmin) +
rmin (hdfio2.cpp:1234-1256).  bits 0 writes float32 with deflate.

The reference ships no HDF5 files, so this module is pinned by round trips
and by the closed forms of the bit reduction (parity unpinned against
EMAN2-written files).
"""
import ctypes
import math
import os

import numpy as np

_LIB = None
_T = {}

H5F_ACC_RDONLY, H5F_ACC_RDWR, H5F_ACC_TRUNC = 0, 1, 2
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0
H5T_INTEGER, H5T_FLOAT, H5T_STRING, H5T_COMPOUND = 0, 1, 3, 6
_hid = ctypes.c_int64
_hsize = ctypes.c_ulonglong
_herr = ctypes.c_int

_SIGS = {
    "H5open": (_herr, []), "H5Eset_auto2": (_herr, [_hid, ctypes.c_void_p, ctypes.c_void_p]),
    "H5Fcreate": (_hid, [ctypes.c_char_p, ctypes.c_uint, _hid, _hid]),
    "H5Fopen": (_hid, [ctypes.c_char_p, ctypes.c_uint, _hid]), "H5Fclose": (_herr, [_hid]),
    "H5Gcreate2": (_hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid]), "H5Gopen2": (_hid, [_hid, ctypes.c_char_p, _hid]),
    "H5Gclose": (_herr, [_hid]),
    "H5Screate_simple": (_hid, [ctypes.c_int, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
    "H5Screate": (_hid, [ctypes.c_int]), "H5Sclose": (_herr, [_hid]),
    "H5Sget_simple_extent_ndims": (ctypes.c_int, [_hid]),
    "H5Sget_simple_extent_dims": (ctypes.c_int, [_hid, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
    "H5Sget_simple_extent_npoints": (ctypes.c_longlong, [_hid]),
    "H5Pcreate": (_hid, [_hid]), "H5Pset_chunk": (_herr, [_hid, ctypes.c_int, ctypes.POINTER(_hsize)]),
    "H5Pset_deflate": (_herr, [_hid, ctypes.c_uint]), "H5Pclose": (_herr, [_hid]),
    "H5Dcreate2": (_hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid, _hid, _hid]),
    "H5Dopen2": (_hid, [_hid, ctypes.c_char_p, _hid]),
    "H5Dwrite": (_herr, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
    "H5Dread": (_herr, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
    "H5Dget_space": (_hid, [_hid]), "H5Dget_type": (_hid, [_hid]), "H5Dclose": (_herr, [_hid]),
    "H5Dget_create_plist": (_hid, [_hid]), "H5Pget_layout": (ctypes.c_int, [_hid]),
    "H5Tget_class": (ctypes.c_int, [_hid]), "H5Tget_size": (ctypes.c_size_t, [_hid]),
    "H5Tget_sign": (ctypes.c_int, [_hid]), "H5Tclose": (_herr, [_hid]), "H5Tcopy": (_hid, [_hid]),
    "H5Tset_size": (_herr, [_hid, ctypes.c_size_t]), "H5Tcreate": (_hid, [ctypes.c_int, ctypes.c_size_t]),
    "H5Tinsert": (_herr, [_hid, ctypes.c_char_p, ctypes.c_size_t, _hid]), "H5Tpack": (_herr, [_hid]),
    "H5Tget_nmembers": (ctypes.c_int, [_hid]),
    "H5Acreate2": (_hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid, _hid]),
    "H5Awrite": (_herr, [_hid, _hid, ctypes.c_void_p]), "H5Aread": (_herr, [_hid, _hid, ctypes.c_void_p]),
    "H5Aopen": (_hid, [_hid, ctypes.c_char_p, _hid]), "H5Aexists": (ctypes.c_int, [_hid, ctypes.c_char_p]),
    "H5Adelete": (_herr, [_hid, ctypes.c_char_p]), "H5Aclose": (_herr, [_hid]),
    "H5Aget_type": (_hid, [_hid]), "H5Aget_space": (_hid, [_hid]),
    "H5Aget_num_attrs": (ctypes.c_int, [_hid]),
    "H5Aopen_by_idx": (_hid, [_hid, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _hsize, _hid, _hid]),
    "H5Aget_name": (ctypes.c_ssize_t, [_hid, ctypes.c_size_t, ctypes.c_char_p]),
    "H5Lexists": (ctypes.c_int, [_hid, ctypes.c_char_p, _hid]),
}


def _lib():
    """the HDF5 C library (HDF5_LIB, else the usual install locations)"""
    global _LIB
    if _LIB is not None:
        return _LIB
    cands = [os.environ.get("HDF5_LIB"), "/opt/conda/lib/libhdf5.so", "libhdf5.so",
             "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so"]
    err = None
    for c in cands:
        if not c:
            continue
        try:
            L = ctypes.CDLL(c)
            break
        except OSError as e:
            err = e
    else:
        raise OSError("libhdf5 not found (%s)" % err)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype, fn.argtypes = res, args
    L.H5open()
    L.H5Eset_auto2(0, None, None)      # no console error stack (hdfio2.cpp:443)
    for k in ("NATIVE_FLOAT", "NATIVE_DOUBLE", "NATIVE_INT", "NATIVE_UINT", "NATIVE_SCHAR", "NATIVE_UCHAR",
              "NATIVE_USHORT", "NATIVE_SHORT", "C_S1"):
        _T[k] = _hid.in_dll(L, "H5T_%s_g" % k).value
    _T["DATASET_CREATE"] = _hid.in_dll(L, "H5P_CLS_DATASET_CREATE_ID_g").value
    _LIB = L
    return L


def _check(v, what):
    if v < 0:
        raise IOError("HDF5: %s failed" % what)
    return v


def _dims(shape):
    return (_hsize * len(shape))(*shape)


def _transform_type():
    L = _lib()
    t = L.H5Tcreate(H5T_COMPOUND, 48)
    for k, name in enumerate(("00", "01", "02", "03", "10", "11", "12", "13", "20", "21", "22", "23")):
        L.H5Tinsert(t, name.encode(), 4 * k, _T["NATIVE_FLOAT"])
    L.H5Tpack(t)
    return t


def _write_attr(loc, name, value):
    """HdfIO2::write_attr (hdfio2.cpp:234-425) for the Python types a header holds"""
    L = _lib()
    bname = name.encode()
    if L.H5Aexists(loc, bname) > 0:
        L.H5Adelete(loc, bname)
    buf = None
    if isinstance(value, (bool, np.bool_)):
        typ, spc = L.H5Tcopy(_T["NATIVE_SCHAR"]), L.H5Screate_simple(1, _dims((1,)), None)
        buf = ctypes.c_char(b"T" if value else b"F")
    elif isinstance(value, (int, np.integer)):
        typ, spc = L.H5Tcopy(_T["NATIVE_INT"]), L.H5Screate_simple(1, _dims((1,)), None)
        buf = ctypes.c_int(int(value))
    elif isinstance(value, (float, np.floating)) and not isinstance(value, np.float64):
        typ, spc = L.H5Tcopy(_T["NATIVE_FLOAT"]), L.H5Screate_simple(1, _dims((1,)), None)
        buf = ctypes.c_float(float(value))
    elif isinstance(value, np.float64):
        typ, spc = L.H5Tcopy(_T["NATIVE_DOUBLE"]), L.H5Screate_simple(1, _dims((1,)), None)
        buf = ctypes.c_double(float(value))
    elif isinstance(value, str):
        b = value.encode()
        typ = L.H5Tcopy(_T["C_S1"])
        L.H5Tset_size(typ, len(b) + 1)
        spc = L.H5Screate(H5S_SCALAR)
        buf = ctypes.create_string_buffer(b, len(b) + 1)
    else:
        a = np.asarray(value)
        if a.shape == (3, 4) and a.dtype.kind == "f":
            typ, spc = _transform_type(), L.H5Screate_simple(1, _dims((1,)), None)
            buf = np.ascontiguousarray(a, np.float32)
        elif a.ndim == 1 and a.dtype.kind in "iu":
            typ, spc = L.H5Tcopy(_T["NATIVE_INT"]), L.H5Screate_simple(1, _dims((a.size,)), None)
            buf = np.ascontiguousarray(a, np.int32)
        elif a.ndim == 1 and a.dtype.kind == "f":
            typ, spc = L.H5Tcopy(_T["NATIVE_FLOAT"]), L.H5Screate_simple(1, _dims((a.size,)), None)
            buf = np.ascontiguousarray(a, np.float32)
        else:
            raise TypeError("HDF5 attribute %s: unsupported value %r" % (name, value))
    attr = _check(L.H5Acreate2(loc, bname, typ, spc, H5P_DEFAULT, H5P_DEFAULT), "H5Acreate " + name)
    ptr = buf.ctypes.data if isinstance(buf, np.ndarray) else ctypes.addressof(buf)
    _check(L.H5Awrite(attr, typ, ctypes.c_void_p(ptr)), "H5Awrite " + name)
    L.H5Aclose(attr)
    L.H5Tclose(typ)
    L.H5Sclose(spc)


def _read_attr(attr):
    """the value of an open attribute, by its stored type"""
    L = _lib()
    typ = L.H5Aget_type(attr)
    spc = L.H5Aget_space(attr)
    n = max(1, int(L.H5Sget_simple_extent_npoints(spc)))
    cls, size = L.H5Tget_class(typ), int(L.H5Tget_size(typ))
    try:
        if cls == H5T_STRING:
            buf = ctypes.create_string_buffer(size * n + 1)
            _check(L.H5Aread(attr, typ, buf), "H5Aread")
            return buf.value.decode(errors="replace")
        if cls == H5T_COMPOUND and size == 48 * 1:
            out = np.zeros(12, np.float32)
            mt = _transform_type()
            _check(L.H5Aread(attr, mt, ctypes.c_void_p(out.ctypes.data)), "H5Aread")
            L.H5Tclose(mt)
            return out.reshape(3, 4)
        if cls == H5T_INTEGER and size == 1:
            out = np.zeros(n, np.int8)
            _check(L.H5Aread(attr, _T["NATIVE_SCHAR"], ctypes.c_void_p(out.ctypes.data)), "H5Aread")
            return bytes([int(out[0]) & 0xFF]) == b"T"
        if cls == H5T_INTEGER:
            unsigned = L.H5Tget_sign(typ) == 0
            out = np.zeros(n, np.uint32 if unsigned else np.int32)
            _check(L.H5Aread(attr, _T["NATIVE_UINT" if unsigned else "NATIVE_INT"], ctypes.c_void_p(out.ctypes.data)),
                   "H5Aread")
            return int(out[0]) if n == 1 else out
        if cls == H5T_FLOAT:
            dbl = size == 8
            out = np.zeros(n, np.float64 if dbl else np.float32)
            _check(L.H5Aread(attr, _T["NATIVE_DOUBLE" if dbl else "NATIVE_FLOAT"], ctypes.c_void_p(out.ctypes.data)),
                   "H5Aread")
            return (float(out[0]) if n == 1 else out)
        return None
    finally:
        L.H5Tclose(typ)
        L.H5Sclose(spc)


def _read_group_attrs(grp, prefix="EMAN."):
    L = _lib()
    out = {}
    for i in range(L.H5Aget_num_attrs(grp)):
        a = L.H5Aopen_by_idx(grp, b".", 0, 0, i, H5P_DEFAULT, H5P_DEFAULT)
        if a < 0:
            continue
        nb = ctypes.create_string_buffer(256)
        L.H5Aget_name(a, 255, nb)
        name = nb.value.decode()
        val = _read_attr(a)
        L.H5Aclose(a)
        if name.startswith(prefix):
            out[name[len(prefix):]] = val
        elif not prefix:
            out[name] = val
    return out


# ---------------------------------------------------------------- bit reduction
def render_limits(data, bits, rmin=None, rmax=None):
    """EMUtil::getRenderMinMax (libEM/emutil.cpp:1739-1860): the storage range
    of a bit-reduced image from its clamped statistics; rmin / rmax are the
    caller's limits (render_min / render_max), the full range when absent"""
    d = np.asarray(data, np.float32).ravel()
    if bits <= 0:
        return 0.0, 0.0, bits
    bits = min(int(bits), 16)
    lo = np.float32(d.min()) if rmin is None else np.float32(rmin)
    hi = np.float32(d.max()) if rmax is None else np.float32(rmax)
    haslim = not (hi <= lo or math.isnan(lo) or math.isnan(hi) or abs(lo) > 1e28 or abs(hi) > 1e28)
    v = np.clip(d, lo, hi) if haslim else d
    mn, mx = np.float32(v.min()), np.float32(v.max())
    nint = int(np.count_nonzero(v == np.floor(v).astype(np.float32)))
    bitval = 1 << bits
    rmin_o, rmax_o = lo, hi          # passed through unless a branch below sets them
    if np.float32(mx - mn) == 1:
        rmin_o, rmax_o = mn, mx
    elif nint == v.size:
        if mx - mn < bitval:
            rmin_o, rmax_o = mn, np.float32(mn + bitval - 1)
        else:
            needed = int(math.ceil(math.log(float(mx - mn) + 1) / math.log(2.0)))
            step = 1 << (needed - bits)
            rmin_o = np.float32(round(float(mn) / step) * step) if (mn < 0 < mx) else mn
            rmax_o = np.float32(rmin_o + step * (bitval - 1))
    elif mn < 0 < mx:
        step = np.float32(np.float32(mx - mn) / np.float32(bitval - 1))
        if mn == np.float32(math.floor(float(mn / step)) * step):
            rmin_o, rmax_o = mn, mx
        else:
            step = np.float32(np.float32(mx - mn) / np.float32(bitval - 2))
            rmin_o = np.float32(math.floor(float(mn / step)) * step)
            rmax_o = np.float32(rmin_o + step * np.float32(bitval - 1))
    return float(rmin_o), float(rmax_o), bits


def render(data, rmin, rmax, bits):
    """Renderer::getRenderedDataAndRendertrunc (renderer.h:74-110) for the
    unsigned types: (integers, count of clamped values)"""
    d = np.asarray(data, np.float32)
    RMIN, RMAX = np.float32(0.0), np.float32((1 << bits) - 1)
    rmin, rmax = np.float32(rmin), np.float32(rmax)
    if rmax == rmin:
        # a constant image: 0/0 in the reference, whose integer cast gives 0
        return np.zeros(d.shape, np.uint8 if bits <= 8 else np.uint16), 0
    x = (d - rmin) / (rmax - rmin) * (RMAX - RMIN) + RMIN     # float32, as the reference
    scaled = np.where(x >= 0, np.floor(x.astype(np.float64) + 0.5), np.ceil(x.astype(np.float64) - 0.5))  # roundf
    out = np.where(d < rmin, RMIN, np.where(d > rmax, RMAX, scaled))
    trunc = int(np.count_nonzero((d < rmin) | (d > rmax)))
    return out.astype(np.uint8 if bits <= 8 else np.uint16), trunc


def unrender(ints, rmin, rmax, bits):
    """the rescale on read (hdfio2.cpp:1251-1252)"""
    RUMAX = np.float32((1 << bits) - 1.0)
    return ((ints.astype(np.float32) / RUMAX) * (np.float32(rmax) - np.float32(rmin)) + np.float32(rmin)).astype(
        np.float32)


def outlier_limits(data):
    """write_compressed's nooutliers range (libpyEM/EMAN2.py:3226-3241): a
    4096-bin histogram over [min, max] (EMData::calc_hist, emdata.cpp:
    2428-2458), at most max(n/20000, 8) pixels cut at each end"""
    d = np.asarray(data, np.float32).ravel()
    n = d.size
    maxout = max(n // 20000, 8)
    h0, h1 = float(d.min()), float(d.max())
    if h1 == h0:
        return h0, h1
    hs = (h1 - h0) / 4095
    w = np.float32(4095) / np.float32(h1 - h0)
    v = (d - np.float32(h0)) * w
    j = np.where(v < 0, (v - np.float32(0.5)).astype(np.int64), (v + np.float32(0.5)).astype(np.int64))
    hist = np.bincount(np.clip(j, 0, 4095), minlength=4096)
    sl = 0
    for sl in range(2048):
        if hist[:sl + 1].sum() > maxout:
            break
    sh = 4095
    for sh in range(4095, 2048, -1):
        if hist[sh:].sum() > maxout:
            break
    return sl * hs + h0, sh * hs + h0


# ---------------------------------------------------------------- files
def write_images(path, images, headers=None, bits=None, level=1, render_range=None, erase=True):
    """Write images (a list of 1-, 2- or 3-D arrays, or a stack (n, ny, nx)
    of 2-D images) as /MDF/images/0 .. n-1.  bits None: float32, contiguous;
    0: float32, deflated; 1-16: bit-reduced with the storage range of
    render_limits from render_range (per image (min, max) or None for the
    full range)."""
    L = _lib()
    if isinstance(images, np.ndarray) and images.ndim == 3:
        images = list(images)
    imgs = [np.ascontiguousarray(a, np.float32) for a in images]
    headers = headers or [{} for _ in imgs]
    if erase and os.path.exists(path):
        os.unlink(path)
    f = _check(L.H5Fcreate(path.encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "H5Fcreate")
    try:
        g = _check(L.H5Gcreate2(f, b"/MDF", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Gcreate /MDF")
        L.H5Gclose(g)
        grp = _check(L.H5Gcreate2(f, b"/MDF/images", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Gcreate")
        _write_attr(grp, "imageid_max", len(imgs) - 1)
        for i, (a, hd) in enumerate(zip(imgs, headers)):
            nz, ny, nx = (a.shape if a.ndim == 3 else (1,) + a.shape if a.ndim == 2 else (1, 1) + a.shape)
            ig = _check(L.H5Gcreate2(f, ("/MDF/images/%d" % i).encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                        "H5Gcreate image")
            h = dict(hd)
            h.setdefault("nx", int(nx))
            h.setdefault("ny", int(ny))
            h.setdefault("nz", int(nz))
            for k, v in h.items():
                if k in ("stored_rendermin", "stored_rendermax", "stored_renderbits", "render_min", "render_max",
                         "render_bits"):
                    continue
                _write_attr(ig, "EMAN." + k, v)
            spc = L.H5Screate_simple(a.ndim, _dims(a.shape), None)
            plist = L.H5Pcreate(_T["DATASET_CREATE"])
            data, htype = a, _T["NATIVE_FLOAT"]
            stored = None
            if bits is not None:
                chunk = (8, nx) if a.ndim == 2 else (1, 8, nx) if a.ndim == 3 else (nx,)
                chunk = tuple(min(c, s) for c, s in zip(chunk, a.shape))
                L.H5Pset_chunk(plist, a.ndim, _dims(chunk))
                L.H5Pset_deflate(plist, int(level))
                if bits > 0:
                    rr = render_range[i] if render_range is not None else None
                    rmin, rmax, b = render_limits(a, bits, *(rr if rr else (None, None)))
                    data, trunc = render(a, rmin, rmax, b)
                    htype = _T["NATIVE_UCHAR"] if b <= 8 else _T["NATIVE_USHORT"]
                    stored = (np.float32(rmin), np.float32(rmax), int(b), int(trunc))
            ds = _check(L.H5Dcreate2(f, ("/MDF/images/%d/image" % i).encode(), htype, spc, H5P_DEFAULT, plist,
                                     H5P_DEFAULT), "H5Dcreate")
            data = np.ascontiguousarray(data)
            _check(L.H5Dwrite(ds, htype, H5S_ALL, H5S_ALL, H5P_DEFAULT, ctypes.c_void_p(data.ctypes.data)),
                   "H5Dwrite")
            L.H5Dclose(ds)
            L.H5Pclose(plist)
            L.H5Sclose(spc)
            if stored is not None:
                _write_attr(ig, "EMAN.stored_rendermin", stored[0])
                _write_attr(ig, "EMAN.stored_rendermax", stored[1])
                _write_attr(ig, "EMAN.stored_renderbits", stored[2])
                _write_attr(ig, "EMAN.stored_truncated", stored[3])
            L.H5Gclose(ig)
        L.H5Gclose(grp)
    finally:
        L.H5Fclose(f)


def write_compressed(path, vol, bits=10, nooutliers=True, level=1, header=None):
    """EMData.write_compressed for one image (libpyEM/EMAN2.py:3142-3258; the
    e2make3dpar.py:494 map output): bits < 0 plain float, 0 float deflated,
    else bit-reduced over the nooutliers range (or the full range)"""
    a = np.asarray(vol, np.float32)
    if bits < 0:
        return write_images(path, [a], [header or {}])
    hd = dict(header or {})
    hd["render_compress_level"] = int(level)
    rr = [outlier_limits(a) if (nooutliers and bits > 0) else None]
    return write_images(path, [a], [hd], bits=bits, level=level, render_range=rr)


class HdfStack:
    """an EMAN2 HDF5 file's images: len(), header(i), read_image(i)"""

    def __init__(self, path):
        L = _lib()
        self.path = path
        self._f = _check(L.H5Fopen(path.encode(), H5F_ACC_RDONLY, H5P_DEFAULT), "H5Fopen " + path)
        grp = L.H5Gopen2(self._f, b"/MDF/images", H5P_DEFAULT)
        if grp < 0:
            L.H5Fclose(self._f)
            raise IOError("%s: HDF5 file has no image data (no /MDF group)" % path)
        a = L.H5Aopen(grp, b"imageid_max", H5P_DEFAULT)
        self.n = int(_read_attr(a)) + 1 if a >= 0 else 0
        if a >= 0:
            L.H5Aclose(a)
        L.H5Gclose(grp)

    def __len__(self):
        return self.n

    def close(self):
        if self._f is not None:
            _lib().H5Fclose(self._f)
            self._f = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def header(self, i):
        L = _lib()
        g = _check(L.H5Gopen2(self._f, ("/MDF/images/%d" % i).encode(), H5P_DEFAULT), "image %d" % i)
        try:
            return _read_group_attrs(g)
        finally:
            L.H5Gclose(g)

    def read_image(self, i):
        L = _lib()
        ds = _check(L.H5Dopen2(self._f, ("/MDF/images/%d/image" % i).encode(), H5P_DEFAULT), "image %d" % i)
        try:
            spc = L.H5Dget_space(ds)
            nd = L.H5Sget_simple_extent_ndims(spc)
            dims = (_hsize * nd)()
            L.H5Sget_simple_extent_dims(spc, dims, None)
            L.H5Sclose(spc)
            shape = tuple(int(d) for d in dims)
            typ = L.H5Dget_type(ds)
            cls, size = L.H5Tget_class(typ), int(L.H5Tget_size(typ))
            L.H5Tclose(typ)
            if cls == H5T_FLOAT:
                out, mt = np.zeros(shape, np.float32), _T["NATIVE_FLOAT"]
            elif size == 1:
                out, mt = np.zeros(shape, np.uint8), _T["NATIVE_UCHAR"]
            elif size == 2:
                out, mt = np.zeros(shape, np.uint16), _T["NATIVE_USHORT"]
            else:
                raise IOError("%s: image %d has a data type EMAN does not support" % (self.path, i))
            _check(L.H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, ctypes.c_void_p(out.ctypes.data)), "H5Dread")
        finally:
            L.H5Dclose(ds)
        h = self.header(i)
        b = h.get("stored_renderbits")
        if b is not None and int(b) > 0 and "stored_rendermin" in h and "stored_rendermax" in h:
            return unrender(out, h["stored_rendermin"], h["stored_rendermax"], int(b))
        return out.astype(np.float32)


def read_stack(path):
    """every image of an EMAN2 HDF5 file: (images (n, ...) float32, headers)"""
    s = HdfStack(path)
    try:
        imgs = [s.read_image(i) for i in range(len(s))]
        hds = [s.header(i) for i in range(len(s))]
    finally:
        s.close()
    return (np.stack(imgs) if imgs else np.zeros((0,), np.float32)), hds
