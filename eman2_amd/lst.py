"""EMAN2 particle lists: the fixed-record "#LSX" files of LSXFile
(libpyEM/EMAN2.py:2471-2640) and libEM/io/lstfastio.cpp, plus the older
variable-length "#LST" form, read.

An #LSX file is three header lines -- "#LSX", one comment line, "# <L>" with
L the record length including its newline -- then records of exactly L
bytes: "<index in the image file>\\t<image file>[\\t<JSON dict>]" padded with
spaces.  The JSON dict overrides the particle's header; a Transform is
{"__class__": "Transform", "matrix": "[m00,m01,...,m23]"}, the twelve 3x4
values printed with %1.6g (EMAN2jsondb.py:192-206, transform.cpp:169-178).

Particles listed here are read from MRC stacks (one particle per section)
or EMAN2 HDF5 stacks (eman2_amd/hdf.py);
other image formats are not read.
"""
import json
import os

import numpy as np

MAGIC = "#LSX"
DEFAULT_COMMENT = ("# This file is in fast LST format. All lines after the next line have exactly the number of "
                   "characters shown on the next line. This MUST be preserved if editing.")


def transform_to_json(m):
    """{"__class__": "Transform", "matrix": "[...]"} of a 3x4 matrix, the
    reference's get_matrix_string(6) text"""
    m = np.asarray(m, np.float64).reshape(3, 4)
    return {"__class__": "Transform", "matrix": "[" + ",".join("%1.6g" % v for v in m.ravel()) + "]"}


def _decode(obj):
    if isinstance(obj, dict) and obj.get("__class__") == "Transform":
        mat = obj["matrix"]
        if isinstance(mat, str):
            mat = [float(v) for v in mat[1:-1].split(",")]
        return np.asarray(mat, np.float32).reshape(3, 4)
    return obj


def _encode(obj):
    if isinstance(obj, np.ndarray) and obj.size == 12:
        return transform_to_json(obj)
    if hasattr(obj, "matrix") and np.asarray(obj.matrix).size == 12:  # a Transform object
        return transform_to_json(obj.matrix)
    raise TypeError("cannot store %r in a particle list" % type(obj))


def _parse(line):
    parts = line.strip().split("\t")
    if len(parts) < 2:
        raise ValueError("particle list record %r has no image file" % line)
    d = {}
    if len(parts) > 2 and len(parts[2]) >= 2:
        d = json.loads(parts[2], object_hook=_decode)
    return int(parts[0]), parts[1], d


def read(path):
    """[(index, image file, dict)] of every record; Transforms in the dicts
    come back as float32 3x4 matrices"""
    with open(path, "rb") as f:
        raw = f.read()
    if b"\r" in raw:
        raise ValueError("%s: contains \\r (a list edited with Windows line ends)" % path)
    text = raw.decode()
    lines = text.split("\n")
    if lines[0] == MAGIC:
        try:
            L = int(lines[2][1:])
        except (IndexError, ValueError):
            raise ValueError("%s: invalid record length line" % path)
        base = len(lines[0]) + len(lines[1]) + len(lines[2]) + 3
        body = text[base:]
        # the record count is the whole records after the header; a partial tail is not one
        # (lstfastio.cpp:104-105)
        return [_parse(body[i * L:(i + 1) * L]) for i in range(len(body) // L)]
    if lines[0] == "#LST":
        return [_parse(ln) for ln in lines[1:] if ln.strip() and not ln.startswith("#")]
    raise ValueError("%s: not an #LSX or #LST file" % path)


def write(path, records, comment=None):
    """Writes [(index, image file, dict or None)] as an #LSX file; the record
    length is the longest record + 1 + 4, as LSXFile sizes a rewritten file"""
    out = []
    for rec in records:
        n, fname = int(rec[0]), str(rec[1])
        d = rec[2] if len(rec) > 2 else None
        if d:
            js = json.dumps(d, indent=None, sort_keys=True, separators=(",", ":"), default=_encode)
            out.append("%d\t%s\t%s" % (n, fname, js))
        else:
            out.append("%d\t%s" % (n, fname))
    L = max((len(s) for s in out), default=15) + 1 + 4
    comment = DEFAULT_COMMENT if comment is None else comment
    if not comment.startswith("#"):
        comment = "# " + comment
    with open(path, "w", newline="\n") as f:
        f.write("%s\n%s\n# %d\n" % (MAGIC, comment.replace("\n", " "), L))
        for s in out:
            f.write(s.ljust(L - 1) + "\n")


def resolve_image_path(lst_path, fname):
    """Where a record's image file is: as written (absolute, or relative to
    the current directory, which is how the reference opens it,
    libEM/io/lstfastio.cpp:154-164 -- in an EMAN2 project sets/x.lst names
    particles/y.hdf relative to the project root), else relative to the
    list's own folder."""
    if os.path.isabs(fname) or os.path.exists(fname):
        return fname
    alt = os.path.join(os.path.dirname(os.path.abspath(lst_path)), fname)
    return alt if os.path.exists(alt) else fname


def load_particles(path, key="xform.projection"):
    """(images (n, ny, nx) float32, 3x4 matrices (n, 3, 4), records): every
    listed particle read from its MRC stack (paths resolved by
    resolve_image_path), and its Transform under key"""
    from . import mrc
    recs = read(path)
    stacks = {}
    imgs, xfs = [], []
    for n, fname, d in recs:
        p = resolve_image_path(path, fname)
        if p.lower().endswith((".hdf", ".h5")):
            if p not in stacks:
                from . import hdf
                stacks[p] = hdf.HdfStack(p)
            st = stacks[p]
            if not 0 <= n < len(st):
                raise IndexError("%s: particle %d of %s, which holds %d" % (path, n, fname, len(st)))
            imgs.append(np.asarray(st.read_image(n), np.float32))
        elif p.lower().endswith((".mrc", ".mrcs", ".map")):
            if p not in stacks:
                stacks[p] = mrc.read(p)[0]
            st = stacks[p]
            if not 0 <= n < st.shape[0]:
                raise IndexError("%s: particle %d of %s, which holds %d" % (path, n, fname, st.shape[0]))
            imgs.append(np.asarray(st[n], np.float32))
        else:
            raise ValueError("%s: MRC and HDF5 stacks are read (record %d names %s)" % (path, n, fname))
        if key not in d:
            raise KeyError("%s: record for particle %d of %s has no %s" % (path, n, fname, key))
        xfs.append(np.asarray(d[key], np.float32).reshape(3, 4))
    if not imgs:
        return np.zeros((0, 0, 0), np.float32), np.zeros((0, 3, 4), np.float32), recs
    return np.stack(imgs), np.stack(xfs), recs
