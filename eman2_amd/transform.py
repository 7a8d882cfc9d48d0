"""EMAN2 ``Transform`` (eman Euler convention) for the fourier_hip host API.

The reconstructor consumes a 3x4 float32 matrix exactly as libEM's
``Transform::matrix`` holds it (libEM/transform.h:503).  The eman Euler ->
matrix formula is evaluated in float64 and cast to float32
(libEM/transform.cpp:650-667); composition is an FP32 3x4 product
(libEM/transform.cpp:1384-1395).
"""
import math

import numpy as np

_DEG2RAD = math.pi / 180.0  # EMConsts::deg2rad (libEM/emobject.cpp:55-56)
_ERR_LIMIT = 0.000001         # Transform::ERR_LIMIT (libEM/transform.cpp:57)


def _apply_precision(v, precision=_ERR_LIMIT):
    """Util::apply_precision (libEM/util.h:1309-1314): a value within
    `precision` of an integer becomes that integer."""
    c = float(np.ceil(np.float32(v)))
    f = float(math.floor(v))
    if abs(v - c) < precision:
        return c
    if abs(v - f) < precision:
        return f
    return v


def euler_eman_matrix(az, alt, phi):
    """3x4 float32 matrix of ``Transform({"type":"eman",...})``."""
    azp, altp, phip = az * _DEG2RAD, alt * _DEG2RAD, phi * _DEG2RAD
    ca, sa = math.cos(azp), math.sin(azp)
    cl, sl = math.cos(altp), math.sin(altp)
    cp, sp = math.cos(phip), math.sin(phip)
    m = np.zeros((3, 4), np.float32)
    m[0, 0] = cp * ca - cl * sa * sp
    m[0, 1] = cp * sa + cl * ca * sp
    m[0, 2] = sl * sp
    m[1, 0] = -sp * ca - cl * sa * cp
    m[1, 1] = -sp * sa + cl * ca * cp
    m[1, 2] = sl * cp
    m[2, 0] = sl * sa
    m[2, 1] = -sl * ca
    m[2, 2] = cl
    return m


def euler_eman_matrices(az, alt, phi):
    """Arrays of angles (deg) -> (n,3,4) float32.  Uses the libm scalar
    trig of ``math`` per element so the float32 result is the same as the
    C library's (vectorised numpy trig may differ in the last double ulp)."""
    az, alt, phi = (np.atleast_1d(np.asarray(a, np.float64)) for a in (az, alt, phi))
    out = np.zeros((az.shape[0], 3, 4), np.float32)
    for i in range(az.shape[0]):
        out[i] = euler_eman_matrix(float(az[i]), float(alt[i]), float(phi[i]))
    return out


class Transform:
    """Minimal libEM ``Transform`` carrying a 3x4 float32 matrix.

    ``Transform({"type": "eman", "az": .., "alt": .., "phi": ..})`` mirrors the
    reference constructor (libEM/transform.cpp:138-141, :250-283).  The
    insertion uses the rotation part; a translation, scale or x mirror is
    applied to the real-space image by ``preprocess_slice``
    (libEM/reconstructor.cpp:1026-1035), as in the reference.
    """

    def __init__(self, params=None):
        self.matrix = np.eye(3, 4, dtype=np.float32)
        if params is None:
            return
        if isinstance(params, Transform):
            self.matrix = params.matrix.copy()
            return
        if isinstance(params, np.ndarray):
            m = np.asarray(params, np.float32)
            if m.shape == (3, 3):
                self.matrix[:, :3] = m
            else:
                self.matrix = m.reshape(3, 4).copy()
            return
        p = {k.lower(): v for k, v in dict(params).items()}
        typ = str(p.get("type", "eman")).lower()
        if typ != "eman":
            raise ValueError("only the 'eman' Euler convention is supported, got %r" % typ)
        self.matrix = euler_eman_matrix(float(p.get("az", 0.0)), float(p.get("alt", 0.0)),
                                        float(p.get("phi", 0.0)))
        tx, ty, tz = (float(p.get(k, 0.0)) for k in ("tx", "ty", "tz"))
        self.matrix[:, 3] = (tx, ty, tz)
        if "scale" in p and float(p["scale"]) != 1.0:
            self.matrix[:, :3] *= np.float32(p["scale"])
        if p.get("mirror", False):
            self.matrix[0, :] *= -1

    def get_matrix(self):
        return self.matrix.reshape(12).tolist()

    def get_determinant(self):
        """Transform::get_determinant (libEM/transform.cpp:1310-1322): the
        3x3 determinant in double, stored as float, snapped to an integer
        within ERR_LIMIT (Util::apply_precision)."""
        m = self.matrix
        d = float(m[0, 0]) * (float(m[1, 1]) * float(m[2, 2]) - float(m[2, 1]) * float(m[1, 2]))
        d -= float(m[0, 1]) * (float(m[1, 0]) * float(m[2, 2]) - float(m[2, 0]) * float(m[1, 2]))
        d += float(m[0, 2]) * (float(m[1, 0]) * float(m[2, 1]) - float(m[2, 0]) * float(m[1, 1]))
        return _apply_precision(float(np.float32(d)))

    def get_scale_and_mirror(self):
        """Transform::get_scale_and_mirror (libEM/transform.cpp:1291-1308):
        the x mirror is a negative determinant, the scale its cube root."""
        det = self.get_determinant()
        mirror = det < 0
        if mirror:
            det = -det
        if det != 1:
            scale = float(np.float32(det ** (1.0 / 3.0)))
            if scale - int(scale) < _ERR_LIMIT:
                scale = float(int(scale))
        else:
            scale = 1.0
        return _apply_precision(scale), mirror

    def get_rotation(self, euler_type="eman"):
        """Transform::get_rotation("eman") (libEM/transform.cpp:852-898):
        az, alt, phi in degrees; alt 0 or 180 puts the whole in-plane angle
        into phi."""
        if str(euler_type).lower() != "eman":
            raise ValueError("only the 'eman' Euler convention is supported, got %r" % euler_type)
        scale, mirror = self.get_scale_and_mirror()
        if scale == 0:
            raise ValueError("UnexpectedBehaviorException: The determinant of the Transform is 0")
        m = self.matrix.astype(np.float64)
        cosalt = m[2, 2] / scale
        xs = -1.0 if mirror else 1.0
        if cosalt >= 1:
            az, alt, phi = 0.0, 0.0, math.degrees(math.atan2(xs * m[0, 1], xs * m[0, 0]))
        elif cosalt <= -1:
            az, alt, phi = 0.0, 180.0, math.degrees(math.atan2(-xs * m[0, 1], xs * m[0, 0]))
        else:
            az = math.degrees(math.atan2(scale * m[2, 0], -scale * m[2, 1]))
            if m[2, 2] == 0.0:
                alt = 90.0
            else:
                alt = math.degrees(math.atan(math.sqrt(m[2, 0] * m[2, 0] + m[2, 1] * m[2, 1]) / abs(m[2, 2])))
            if m[2, 2] * scale < 0:
                alt = 180.0 - alt
            phi = math.degrees(math.atan2(xs * m[0, 2], m[1, 2]))
        return {"type": "eman", "az": az, "alt": alt, "phi": phi}

    def is_pure_rotation(self, tol=1e-6):
        m = self.matrix.astype(np.float64)
        if np.any(np.abs(m[:, 3]) > tol):
            return False
        det = np.linalg.det(m[:, :3])
        return abs(det - 1.0) < 1e-5

    def __mul__(self, other):
        """FP32 3x4 composition in the reference's summation order."""
        a, b = self.matrix, other.matrix
        r = np.zeros((3, 4), np.float32)
        for i in range(3):
            for j in range(4):
                r[i, j] = (a[i, 0] * b[0, j] + a[i, 1] * b[1, j]) + a[i, 2] * b[2, j]
            r[i, 3] = r[i, 3] + a[i, 3]
        t = Transform()
        t.matrix = r
        return t

    def __repr__(self):
        return "Transform(%s)" % np.array2string(self.matrix, precision=5)


def get_symmetries(sym):
    """Symmetry3D::get_symmetries (libEM/symmetry.cpp:1240-1246): the
    rotation operators of a symmetry name (c<n>, d<n>, h<n>[:1:daz:tz], tet,
    oct, icos) as an (n, 3, 3) float32 array, from the library's own table.
    Unknown names raise NotExistingObjectException as Factory::get does."""
    import ctypes
    from . import _lib
    L = _lib.load()
    n = L.fh_get_symmetries(str(sym).encode(), None, 0)
    if n < 0:
        raise _lib.FourierHipError(-n, "symmetry %r (c<n>, d<n>, h<n>, tet, oct, icos)" % (sym,))
    out = np.zeros((n, 3, 3), np.float32)
    rc = L.fh_get_symmetries(str(sym).encode(), out.ctypes.data_as(ctypes.c_void_p), n)
    if rc != n:
        raise _lib.FourierHipError(-rc if rc < 0 else 6, "symmetry table of %r" % (sym,))
    return out
