"""Seeded synthetic particle data for the reconstruction workloads.

A phantom is a sum of isotropic 3-D Gaussians.  Its projection along a
``Transform`` T is analytic: a 2-D Gaussian of amplitude A*sigma*sqrt(2*pi)
centred at (T c)_{x,y}, following the projection convention
p(r) = sum_z V(T^-1 r) of libEM/projector.cpp:765-805.  Orientations are
uniform on SO(3) (az, phi ~ U[0,360), alt = acos(U[-1,1])).  The CTF is the
CTF_AMP closed form of libEM/ctf.cpp:1006-1021 applied at box sampling,
noise is white Gaussian at a given SNR, and each particle draws from its own
counter-keyed stream ``default_rng([seed, index])`` so any subset can be
regenerated without the rest.
"""
import math

import numpy as np

from .transform import euler_eman_matrices


def phantom(n_gauss, radius, seed, sigma=(2.0, 4.0), amp=(0.5, 1.0)):
    rng = np.random.default_rng([seed, 0x5048414E])
    # uniform in a ball: direction * r^(1/3)
    d = rng.standard_normal((n_gauss, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = radius * rng.random(n_gauss) ** (1.0 / 3.0)
    centres = d * r[:, None]
    sig = rng.uniform(sigma[0], sigma[1], n_gauss)
    a = rng.uniform(amp[0], amp[1], n_gauss)
    return {"centres": centres, "sigmas": sig, "amps": a}


def symmetric_phantom(subunit, sym_mats):
    """Replicate a subunit phantom by symmetry operators (n,3,4 float32):
    the copy under S_k has centres S_k^T c (a rotation of the map)."""
    cs, ss, aa = [], [], []
    for m in sym_mats:
        R = np.asarray(m, np.float64)[:, :3]
        cs.append(subunit["centres"] @ R)
        ss.append(subunit["sigmas"])
        aa.append(subunit["amps"])
    return {"centres": np.concatenate(cs), "sigmas": np.concatenate(ss), "amps": np.concatenate(aa)}


def orientations(n, seed, start=0):
    """Uniform SO(3) eman Euler angles (deg) for particles start..start+n-1."""
    az = np.empty(n)
    alt = np.empty(n)
    phi = np.empty(n)
    for i in range(n):
        rng = np.random.default_rng([seed, start + i, 1])
        u = rng.random(3)
        az[i] = 360.0 * u[0]
        alt[i] = math.degrees(math.acos(2.0 * u[1] - 1.0))
        phi[i] = 360.0 * u[2]
    return az, alt, phi


def project(ph, mats, box):
    """Analytic projections: (n,box,box) float32 for (n,3,4) matrices."""
    mats = np.asarray(mats, np.float64)
    c = ph["centres"]
    s = ph["sigmas"]
    a = ph["amps"] * s * math.sqrt(2.0 * math.pi)
    px = np.arange(box, dtype=np.float64) - box // 2
    out = np.empty((mats.shape[0], box, box), np.float32)
    for i in range(mats.shape[0]):
        R = mats[i, :, :3]
        cx = c @ R[0]
        cy = c @ R[1]
        gx = np.exp(-((px[None, :] - cx[:, None]) ** 2) / (2 * s[:, None] ** 2))  # (g, box)
        gy = np.exp(-((px[None, :] - cy[:, None]) ** 2) / (2 * s[:, None] ** 2))
        out[i] = (gy.T * a[None, :]) @ gx
    return out


def phantom_volume(ph, n):
    """Sampled phantom on an n^3 grid with the origin at n//2 ([z][y][x])."""
    x = np.arange(n, dtype=np.float64) - n // 2
    vol = np.zeros((n, n, n))
    for (cx, cy, cz), s, a in zip(ph["centres"], ph["sigmas"], ph["amps"]):
        gx = np.exp(-((x - cx) ** 2) / (2 * s * s))
        gy = np.exp(-((x - cy) ** 2) / (2 * s * s))
        gz = np.exp(-((x - cz) ** 2) / (2 * s * s))
        vol += a * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    return vol.astype(np.float32)


def ctf_amp_image(box, apix, defocus, voltage=300.0, cs=2.7, ampcont=10.0, bfactor=80.0):
    """CTF_AMP on the rfft2 grid of a box x box image (float64)."""
    lam = 12.2639 / math.sqrt(voltage * 1000.0 + 0.97845 * voltage * voltage)
    g1 = math.pi / 2.0 * cs * 1.0e7 * lam ** 3
    g2 = math.pi * lam * 10000.0
    acac = math.pi / 2.0 - math.asin(ampcont / 100.0)
    ky = np.fft.fftfreq(box)[:, None]
    kx = np.fft.rfftfreq(box)[None, :]
    s = np.sqrt(kx * kx + ky * ky) / apix
    gam = -g1 * s ** 4 + g2 * defocus * s ** 2
    return np.cos(gam - acac) * np.exp(-(bfactor / 4.0) * s * s)


def particles(ph, n, box, seed, start=0, apix=1.0, snr=None, ctf=None, defocus_range=(0.8, 2.5)):
    """Generate particles start..start+n-1.

    Returns (images float32 (n,box,box), matrices (n,3,4), defocus (n,) or None).
    ``ctf`` is None or a dict with voltage/cs/ampcont/bfactor.
    """
    az, alt, phi = orientations(n, seed, start)
    mats = euler_eman_matrices(az, alt, phi)
    imgs = project(ph, mats, box).astype(np.float64)
    dfs = None
    if ctf is not None:
        dfs = np.empty(n, np.float32)
        for i in range(n):
            rng = np.random.default_rng([seed, start + i, 2])
            dfs[i] = rng.uniform(*defocus_range)
            c = ctf_amp_image(box, apix, float(dfs[i]), ctf.get("voltage", 300.0), ctf.get("cs", 2.7),
                              ctf.get("ampcont", 10.0), ctf.get("bfactor", 80.0))
            imgs[i] = np.fft.irfft2(np.fft.rfft2(imgs[i]) * c, s=(box, box))
    if snr is not None:
        for i in range(n):
            rng = np.random.default_rng([seed, start + i, 3])
            sd = math.sqrt(float(imgs[i].var()) / snr) if imgs[i].var() > 0 else 0.0
            imgs[i] += rng.standard_normal((box, box)) * sd
    return imgs.astype(np.float32), mats, dfs


def particles_device(ph, n, box, seed, apix=1.0, snr=None, ctf=None, defocus_range=(0.8, 2.5),
                     device="cuda", chunk=1024):
    """Device-side generator with the same definition as ``particles`` (not
    bit-identical: torch RNG and FP32 on the GPU).  Returns
    (images torch.float32 (n,box,box) on ``device``, matrices (n,3,4) numpy,
    defocus (n,) numpy or None).  Used to make benchmark-sized inputs that
    already live in HBM."""
    import torch

    az, alt, phi = orientations_fast(n, seed)
    mats = euler_eman_matrices(az, alt, phi)
    rng = np.random.default_rng([seed, 7])
    dfs = rng.uniform(*defocus_range, n).astype(np.float32) if ctf is not None else None
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    C = torch.tensor(ph["centres"], dtype=torch.float32, device=device)            # (G,3)
    sig = torch.tensor(ph["sigmas"], dtype=torch.float32, device=device)          # (G,)
    amp = torch.tensor(ph["amps"] * ph["sigmas"] * math.sqrt(2 * math.pi), dtype=torch.float32, device=device)
    px = torch.arange(box, dtype=torch.float32, device=device) - box // 2
    out = torch.empty((n, box, box), dtype=torch.float32, device=device)
    ctfk = None
    if ctf is not None:
        lam = 12.2639 / math.sqrt(ctf.get("voltage", 300.0) * 1000.0 + 0.97845 * ctf.get("voltage", 300.0) ** 2)
        g1 = math.pi / 2.0 * ctf.get("cs", 2.7) * 1.0e7 * lam ** 3
        g2 = math.pi * lam * 10000.0
        acac = math.pi / 2.0 - math.asin(ctf.get("ampcont", 10.0) / 100.0)
        ky = torch.fft.fftfreq(box, device=device)[:, None]
        kx = torch.fft.rfftfreq(box, device=device)[None, :]
        s2 = (kx * kx + ky * ky) / (apix * apix)
        ctfk = (g1, g2, acac, s2, ctf.get("bfactor", 80.0))
    for c0 in range(0, n, chunk):
        m = min(chunk, n - c0)
        R = torch.tensor(mats[c0:c0 + m, :2, :3], dtype=torch.float32, device=device)   # (m,2,3)
        cxy = torch.matmul(R, C.t().contiguous())                                        # (m,2,G)
        gx = torch.exp(-((px[None, None, :] - cxy[:, 0, :, None]) ** 2) / (2 * sig[None, :, None] ** 2))
        gy = torch.exp(-((px[None, None, :] - cxy[:, 1, :, None]) ** 2) / (2 * sig[None, :, None] ** 2))
        img = torch.bmm((gy * amp[None, :, None]).transpose(1, 2).contiguous(), gx.contiguous())
        if ctfk is not None:
            g1, g2, acac, s2, bf = ctfk
            df = torch.tensor(dfs[c0:c0 + m], dtype=torch.float32, device=device)[:, None, None]
            gam = -g1 * s2[None] ** 2 + g2 * df * s2[None]
            k = torch.cos(gam - acac) * torch.exp(-(bf / 4.0) * s2[None])
            img = torch.fft.irfft2(torch.fft.rfft2(img) * k, s=(box, box))
        if snr is not None:
            sd = torch.sqrt(img.var(dim=(1, 2), keepdim=True) / snr)
            img = img + torch.randn(img.shape, generator=g, device=device) * sd
        out[c0:c0 + m] = img
    return out, mats, dfs


def orientations_fast(n, seed):
    """Uniform SO(3) eman Euler angles from one seeded stream (fast path for
    large n; same distribution as ``orientations``)."""
    rng = np.random.default_rng([seed, 11])
    u = rng.random((n, 3))
    return 360.0 * u[:, 0], np.degrees(np.arccos(2.0 * u[:, 1] - 1.0)), 360.0 * u[:, 2]


def good_size(n):
    """FFT-friendly box sizes (libpyEM/EMAN2.py:660-677 good_size)."""
    k = int(n)
    while True:
        m = k
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1 and k % 2 == 0:
            return k
        k += 1


def gausskernelfix(vol, gauss_width=4.0):
    """math.gausskernelfix: f /= exp(-r^2 * gauss_width / ny^2)
    (libEM/processor.h:4131-4176), origin at n//2."""
    n = vol.shape[0]
    x = np.arange(n, dtype=np.float64) - n // 2
    r2 = x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2
    return (vol / np.exp(-r2 * gauss_width / (n * n))).astype(np.float32)


def clip_center(vol, n):
    """Centre clip of a cube to n^3 (e2make3dpar.py:431-432 Region semantics)."""
    N = vol.shape[0]
    o = (N - n) // 2
    return vol[o:o + n, o:o + n, o:o + n]
