"""Agreement metrics used by the parity tests and the benchmark report.

* ``fsc``: Fourier shell correlation of two real cubic maps, shells of
  width 1 voxel up to Nyquist.  A shell empty in both maps counts as 1 (they
  agree); a shell empty in one map only counts as 0 (one map lost it).
* ``max_rel_err``: max |a-b| / max |b| (error relative to the map's peak;
  a per-voxel ratio is ill-defined for sums that cancel to ~0).
* ``weight_rel_err``: per-voxel relative error of the weight volume, whose
  entries are sums of positive terms so the ratio is well defined.
"""
import numpy as np


def fsc(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = a.shape[0]
    fa = np.fft.rfftn(a)
    fb = np.fft.rfftn(b)
    kz = np.fft.fftfreq(n) * n
    ky = np.fft.fftfreq(n) * n
    kx = np.fft.rfftfreq(n) * n
    r = np.sqrt(kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2)
    shell = np.rint(r).astype(int)
    nsh = n // 2 + 1
    m = shell < nsh
    num = np.bincount(shell[m], (fa * np.conj(fb)).real[m], nsh)
    da = np.bincount(shell[m], (np.abs(fa) ** 2)[m], nsh)
    db = np.bincount(shell[m], (np.abs(fb) ** 2)[m], nsh)
    with np.errstate(invalid="ignore", divide="ignore"):
        c = num / np.sqrt(da * db)
    both_empty = (da == 0) & (db == 0)
    c[both_empty] = 1.0
    c[~np.isfinite(c)] = 0.0
    return c


def max_rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a - b).max())


def weight_rel_err(a, b, floor=1e-6):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.maximum(np.abs(b), floor * np.abs(b).max())
    return float((np.abs(a - b) / den).max())


def corr(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    a = a - a.mean()
    b = b - b.mean()
    return float(a @ b / np.sqrt((a @ a) * (b @ b)))
