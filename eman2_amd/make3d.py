"""The caller loop of e2make3d.py's reconstruct() (programs/e2make3d.py:385-539)
driving a reconstructor in batches: niter passes of insertion; between
passes every particle's agreement with the volume (determine_slice_agreement,
its own contribution removed unless it is excluded), its norm multiplied by
reconstruct_norm, absolute qualities turned into relative ones by a running
average over particles sorted by weight, and a quality cutoff that leaves
the worst particles out of the next pass.  The last pass ends with finish().

The reconstructor is anything with setup() / setup_seed() / insert_batch() /
agreement_batch() / finish() -- the "fourier_hip" reconstructor here.  The
reference multiplies each cached preprocessed slice by its norm; the batched
form scales the raw particle instead, which is the same up to float
rounding (edge-mean subtraction, padding and the FFT are linear).
"""
import math
import os

import numpy as np


AGREE_KEYS = ("reconstruct_norm", "reconstruct_absqual", "reconstruct_absqual_lowres", "reconstruct_qual",
              "reconstruct_weight")


def relative_quality(weights, absqual):
    """e2make3d.py:494-517: particles sorted by (weight, absqual, index); each
    absolute quality over the mean of its sorted neighbours [i-10, i+10);
    weight-0 particles get -1 and stay out of the returned list."""
    n = len(weights)
    order = sorted(range(n), key=lambda i: (float(weights[i]), float(absqual[i]), 0, i))
    qual = np.zeros(n)
    qlist = []
    for k, i in enumerate(order):
        if weights[i] == 0:
            qual[i] = -1
            continue
        sub = [float(absqual[order[j]]) for j in range(max(0, k - 10), min(k + 10, n))]
        mean = sum(sub) / len(sub)
        try:
            q = float(absqual[i]) / mean
        except ZeroDivisionError:
            q = -1
        qual[i] = q
        qlist.append(q)
    return qual, qlist


def quality_cutoff(absqual, qlist, keep, keepsig, keepabs):
    """e2make3d.py:519-539"""
    if keepabs:
        sq = np.array(sorted(float(a) for a in absqual))
        if keepsig:
            return float(sq.mean() - sq.std() * keepabs)
        return float(sq[-int(keep * len(qlist))])
    if keepsig:
        qmean = sum(qlist) / len(qlist)
        qsigma = np.sqrt(sum(q * q for q in qlist) / len(qlist) - qmean ** 2)
        return float(qmean - qsigma * keep)
    ql = sorted(qlist)
    return float(ql[-int(keep * len(ql))])


def reconstruct(recon, imgs, xforms, weights=None, niter=2, keep=1.0, keepsig=False, keepabs=False,
                start=None, startweight=0.0, apix=1.0, defocus=None, ctf=None, progress=None):
    """e2make3d.py reconstruct() over numpy particles (n, box, box) and their
    3x4 Transforms.  Returns (map, info): the finished map and per-particle
    norm / qualities plus the included / excluded indices of the last pass
    and ptcl_repr (the summed weight of the inserted particles).
    progress(fraction) is called after each pass (E2progress)."""
    imgs = np.ascontiguousarray(imgs, np.float32)
    n = len(imgs)
    xf = np.ascontiguousarray(xforms, np.float32).reshape(n, 3, 4)
    w = np.ones(n, np.float32) if weights is None else np.ascontiguousarray(weights, np.float32)
    norm = np.ones(n, np.float64)
    absqual = np.zeros(n)
    qual = np.zeros(n)
    qcutoff = None
    out = None
    included, excluded = [], []
    ptcl = 0.0
    for it in range(niter):
        if start is not None:
            recon.setup_seed(start, startweight)
        else:
            recon.setup()
        # the quality test is skipped (try/except) until a cutoff exists
        if qcutoff is None:
            take = np.ones(n, bool)
        else:
            q = absqual if keepabs else qual
            take = ~((q < qcutoff) | (w == 0))
        if it == niter - 1:
            included = [int(i) for i in np.nonzero(take)[0]]
            excluded = [int(i) for i in np.nonzero(~take)[0]]
        idx = np.nonzero(take & (w != 0))[0]
        ptcl = float(w[idx].sum())
        scaled = imgs[idx] * norm[idx].astype(np.float32)[:, None, None]
        if len(idx):
            recon.insert_batch(scaled, xf[idx], w[idx],
                               defocus=None if defocus is None else np.asarray(defocus, np.float32)[idx], ctf=ctf)
        if it != niter - 1:
            # agreement of every particle; its own slice is removed unless it is excluded
            dosub = take & (w != 0)
            res = np.zeros((n, 5), np.float32)
            scaled_all = imgs * norm.astype(np.float32)[:, None, None]
            for flag in (True, False):
                sel = np.nonzero(dosub == flag)[0]
                if len(sel) == 0:
                    continue
                res[sel] = recon.agreement_batch(scaled_all[sel], xf[sel], w[sel],
                                                 defocus=None if defocus is None else
                                                 np.asarray(defocus, np.float32)[sel],
                                                 ctf=ctf, sub=flag, apix=apix)
            norm *= res[:, 0]
            absqual = res[:, 1].astype(np.float64)
            qual, qlist = relative_quality(w, absqual)
            if qlist:
                qcutoff = quality_cutoff(absqual, qlist, keep, keepsig, keepabs)
        else:
            out = recon.finish(True)
        if progress is not None:
            progress((it + 1) / niter)
    info = dict(norm=norm, reconstruct_absqual=absqual, reconstruct_qual=qual, qcutoff=qcutoff,
                threed_ptcl_idxs=included, threed_excl_ptcl_idxs=excluded, ptcl_repr=ptcl)
    return out, info


def fill_orientations(xform, fillangle, pad):
    """e2make3dpar.py --fillangle (reconstruct(), :666-677, 737-747): the
    orientations a particle is inserted at and their weight factors.  The
    step is atan2(1, P/2) degrees (P = max(pad)); fillangle / step is cut
    to at most 9 and at most 1 means no filling (the particle's own
    Transform, factor 1); otherwise a den x den grid of (alt, az) offsets
    from -fillangle/2 with step fillangle/(den-1) - 1e-5, each the eman
    rotation (az + daz, alt + dalt, phi 0) with factor
    exp(-(dalt^2 + daz^2) / (fillangle/4)^2).  Returns [(3x4 float32, factor)]:
    the fill rotations carry the particle's own scale, x mirror and
    translation, which preprocess_slice applies to the image before the
    insertion, as the reference preprocesses it with the particle's
    Transform and inserts at the new ones."""
    from .transform import Transform, euler_eman_matrix
    m0 = np.asarray(xform, np.float32).reshape(3, 4)
    P = max(pad) if np.ndim(pad) else pad
    astep = math.atan2(1.0, P / 2.0) * 180.0 / math.pi
    den = math.floor(fillangle / astep)
    if den > 9:
        den = 9
    if den <= 1:
        return [(m0.copy(), 1.0)]
    astep = fillangle / (den - 1) - .00001
    t = Transform(m0)
    rot = t.get_rotation("eman")
    scale, mirror = t.get_scale_and_mirror()
    out = []
    for dalt in np.arange(-fillangle / 2.0, fillangle / 2.0, astep):
        for daz in np.arange(-fillangle / 2.0, fillangle / 2.0, astep):
            wm = math.exp(-(dalt ** 2 + daz ** 2) / (fillangle / 4.0) ** 2)
            m = euler_eman_matrix(rot["az"] + daz, rot["alt"] + dalt, 0.0)
            if scale != 1.0:
                m[:, :3] *= np.float32(scale)
            if mirror:
                m[0, :3] *= -1
            m[:, 3] = m0[:, 3]
            out.append((m, wm))
    return out


def insert_with_fill(recon, imgs, xforms, weights=None, fillangle=0.0, pad=None, defocus=None, ctf=None,
                     progress=None):
    """e2make3dpar.py reconstruct()'s insertion with --fillangle (:737-747):
    every particle inserted at each of its fill_orientations() with its
    weight times the factor -- batched as one insert_batch per grid offset
    (every particle has the same den x den grid).  pad defaults to the
    reconstructor's size.  progress(fraction) is called after each batch
    (E2progress)."""
    imgs = np.ascontiguousarray(imgs, np.float32)
    n = len(imgs)
    xf = np.ascontiguousarray(xforms, np.float32).reshape(n, 3, 4)
    w = np.ones(n, np.float32) if weights is None else np.ascontiguousarray(weights, np.float32)
    pad = recon.P if pad is None else pad
    fills = [fill_orientations(xf[i], fillangle, pad) for i in range(n)]
    k = len(fills[0]) if n else 0
    for j in range(k):
        xj = np.stack([f[j][0] for f in fills])
        wj = (w * np.array([f[j][1] for f in fills], np.float64)).astype(np.float32)
        recon.insert_batch(imgs, xj, wj, defocus=defocus, ctf=ctf)
        if progress is not None:
            progress((j + 1) / k)


def threshold_compress(vol):
    """threshold.compress at the map's mean with its standard deviation as
    the range (e2make3dpar.py:386-390; CollapseProcessor, processor.h:
    3274-3317 -- clamponly is declared but not used by process_pixel):
    x > v + r -> x - r, x < v - r -> x + r, else v.  mean and sigma as
    EMData::update_stat (emdata.cpp:3154-3204): double sums, the n - 1
    variance, stored as float."""
    a = np.asarray(vol, np.float32)
    d = a.astype(np.float64).ravel()
    n = d.size
    s1, s2 = d.sum(), (d * d).sum()
    v = np.float32(s1 / n)
    var = (s2 - s1 * s1 / n) / (n - 1)
    r = np.float32(math.sqrt(var) if var >= 0 else 0.0)
    return np.where(a > v + r, a - r, np.where(a < v - r, a + r, v)).astype(np.float32)


def reconstruct_iterative(recon, imgs, xforms, weights=None, niter=4, seed=None, defocus=None, ctf=None,
                          progress=None):
    """e2make3dpar.py --iterative (:241-246, 353-392): niter passes of a
    fourier_iter reconstructor, each inserting every particle and ending
    with finish(True); the first starts from setup() (or the seed map), each
    later one from setup_seed(previous map, 1.0), the previous map first
    compressed about its mean by its standard deviation.  Returns the last
    map (numpy, P^3)."""
    imgs = np.ascontiguousarray(imgs, np.float32)
    n = len(imgs)
    xf = np.ascontiguousarray(xforms, np.float32).reshape(n, 3, 4)
    w = np.ones(n, np.float32) if weights is None else np.ascontiguousarray(weights, np.float32)
    out = None
    for it in range(niter):
        if it == 0:
            if seed is not None:
                recon.setup_seed(seed, 1.0)
            else:
                recon.setup()
        else:
            recon.setup_seed(out, 1.0)
        recon.insert_batch(imgs, xf, w, defocus=defocus, ctf=ctf)
        out = recon.finish(True)
        out = out.numpy() if hasattr(out, "numpy") else np.asarray(out)
        if it < niter - 1:
            out = threshold_compress(out)
        if progress is not None:
            progress((it + 1) / niter)
    return out


def spa_make3d(make_recon, imgs, xforms, weights=None, ntasks=4, defocus=None, ctf=None, progress=None):
    """e2spa_make3d.py's split and merge (:154-192, Make3dTask :351-395):
    particle i goes to task i % ntasks; each task inserts its particles of
    weight > 0 into its own reconstructor (make_recon() -> a reconstructor
    not yet set up) and ends with finish(False) and its normout; the tasks'
    volumes times their normouts and the normouts are summed on a fresh
    reconstructor, the sum multiplied by the reciprocal of the summed
    normouts and inverse-transformed.  Returns the map (P^3, numpy)."""
    imgs = np.ascontiguousarray(imgs, np.float32)
    n = len(imgs)
    xf = np.ascontiguousarray(xforms, np.float32).reshape(n, 3, 4)
    w = np.ones(n, np.float32) if weights is None else np.ascontiguousarray(weights, np.float32)
    parts = []
    for t in range(ntasks):
        idx = np.arange(t, n, ntasks)
        idx = idx[w[idx] > 0]                     # reconstruct() skips wt <= 0 (e2spa_make3d.py:304-306)
        r = make_recon()
        r.setup()
        if len(idx):
            c = ctf
            if ctf is not None and ctf.get("dfdf") is not None:
                c = dict(ctf, dfdf=np.asarray(ctf["dfdf"], np.float32)[idx])
            r.insert_batch(imgs[idx], xf[idx], w[idx],
                           defocus=None if defocus is None else np.asarray(defocus, np.float32)[idx], ctf=c)
        normout = np.zeros((r.P, r.P, r.P // 2 + 1), np.float32)
        r.set_params({"normout": normout})
        parts.append((r.finish(False, return_numpy=True), normout))
        if progress is not None:
            progress((t + 1) / ntasks)
    m = make_recon()
    m.setup()
    for threed, nv in parts:
        m.merge_add(threed, nv)
    return m.merge_finish(return_numpy=True)


def reconstruct_mrc(recon, stack_path, xforms, weights=None, out_path=None, **kw):
    """reconstruct() over a particle stack in an MRC file (memory-mapped,
    one particle per section); the map goes to out_path when given, with
    the stack's pixel size."""
    from . import mrc
    stack, h = mrc.read(stack_path)
    out, info = reconstruct(recon, stack, xforms, weights, apix=kw.pop("apix", h["apix"]), **kw)
    if out_path is not None:
        vol = out.numpy() if hasattr(out, "numpy") else np.asarray(out)
        mrc.write(out_path, vol, apix=h["apix"])
    return out, info


def reconstruct_lst(recon, lst_path, out_path=None, apix=None, **kw):
    """reconstruct() over the particles of an #LSX / #LST list (MRC stacks,
    orientations from each record's xform.projection); the map goes to
    out_path when given"""
    from . import lst, mrc
    imgs, xforms, recs = lst.load_particles(lst_path)
    if apix is None:
        first = lst.resolve_image_path(lst_path, recs[0][1]) if recs else None
        if first is None:
            apix = 1.0
        elif first.lower().endswith((".hdf", ".h5")):
            from . import hdf
            apix = float(hdf.HdfStack(first).header(recs[0][0]).get("apix_x", 1.0))
        else:
            apix = mrc.read_header(first)["apix"]
    out, info = reconstruct(recon, imgs, xforms, apix=apix, **kw)
    if out_path is not None:
        write_map(out_path, out, apix)
    return out, info


def write_map(path, vol, apix, compressbits=10):
    """the finished map to disk: an HDF5 path as e2make3dpar.py:494 writes it
    (write_compressed with compressbits, 10 by default; negative for plain
    float), anything else as MRC2014"""
    from . import mrc
    v = vol.numpy() if hasattr(vol, "numpy") else np.asarray(vol, np.float32)
    if path.lower().endswith((".hdf", ".h5")):
        from . import hdf
        hdf.write_compressed(path, v, bits=compressbits,
                             header={"apix_x": float(apix), "apix_y": float(apix), "apix_z": float(apix)})
    else:
        mrc.write(path, v, apix=apix)


def reconstruct_hdf(recon, stack_path, key="xform.projection", out_path=None, compressbits=10, **kw):
    """reconstruct() over an EMAN2 HDF5 particle stack whose images carry
    their orientation under key (and apix_x); the map goes to out_path when
    given (write_map)"""
    from . import hdf
    imgs, heads = hdf.read_stack(stack_path)
    if any(key not in h for h in heads):
        raise KeyError("%s: an image has no %s" % (stack_path, key))
    xforms = np.stack([np.asarray(h[key], np.float32).reshape(3, 4) for h in heads])
    apix = kw.pop("apix", float(heads[0].get("apix_x", 1.0)) if heads else 1.0)
    out, info = reconstruct(recon, imgs, xforms, apix=apix, **kw)
    if out_path is not None:
        write_map(out_path, out, apix, compressbits)
    return out, info
