"""MRC / MRCS files (the MRC2014 layout EMAN2 reads and writes in
libEM/io/mrcio.cpp): particle stacks in, reconstructed maps out.

A file is a 1024-byte header of 4-byte words -- nx, ny, nz, mode, the
start indices, the sampling mx, my, mz, the cell lengths in Angstrom
(pixel size = cell length / sampling), the axis order, min / max / mean,
the space group, the extended-header length (nsymbt), the origin, the
"MAP " tag, the machine stamp, rms and ten 80-character labels -- then
nsymbt bytes of extended header, then the data with x fastest.  Stacks of
2-D particles put one particle per section (z).
"""
import os
import struct

import numpy as np

_MODES = {0: np.int8, 1: np.int16, 2: np.float32, 6: np.uint16, 12: np.float16}
HEADER_BYTES = 1024


def read_header(path):
    with open(path, "rb") as f:
        raw = f.read(HEADER_BYTES)
    if len(raw) < HEADER_BYTES:
        raise ValueError("%s: shorter than an MRC header" % path)
    # machine stamp 0x44 0x41 / 0x44 0x44: little endian, 0x11 0x11: big endian
    end = ">" if raw[212] == 0x11 else "<"
    w = struct.unpack(end + "10i6f3i3f2i", raw[:96])
    nx, ny, nz, mode = w[0:4]
    mx, my, mz = w[7:10]
    cella = w[10:13]
    h = dict(nx=nx, ny=ny, nz=nz, mode=mode, mx=mx, my=my, mz=mz, cella=cella,
             mapc=w[16], mapr=w[17], maps=w[18], dmin=w[19], dmax=w[20], dmean=w[21],
             ispg=w[22], nsymbt=w[23], endian=end)
    h["origin"] = struct.unpack(end + "3f", raw[196:208])
    h["apix"] = float(cella[0]) / mx if mx else 1.0
    if mode not in _MODES:
        raise ValueError("%s: MRC mode %d is not supported" % (path, mode))
    return h


def read(path, mmap=True):
    """(data, header): data (nz, ny, nx) in the file's type (a read-only
    memory map unless mmap is False); for a particle stack nz is the count."""
    h = read_header(path)
    dt = np.dtype(_MODES[h["mode"]]).newbyteorder(h["endian"])
    off = HEADER_BYTES + h["nsymbt"]
    shape = (h["nz"], h["ny"], h["nx"])
    need = off + int(np.prod(shape)) * dt.itemsize
    if os.path.getsize(path) < need:
        raise ValueError("%s: %d bytes, the header describes %d" % (path, os.path.getsize(path), need))
    if mmap:
        data = np.memmap(path, dtype=dt, mode="r", offset=off, shape=shape)
    else:
        with open(path, "rb") as f:
            f.seek(off)
            data = np.fromfile(f, dtype=dt, count=int(np.prod(shape))).reshape(shape)
    return data, h


def write(path, data, apix=1.0, label="fourier_hip"):
    """Writes a float32 map or stack (nz, ny, nx) (a 2-D array is one section)."""
    a = np.ascontiguousarray(data, np.float32)
    if a.ndim == 2:
        a = a[None]
    nz, ny, nx = a.shape
    hdr = bytearray(HEADER_BYTES)
    mean = float(a.mean()) if a.size else 0.0
    rms = float(a.std()) if a.size else 0.0
    amin = float(a.min()) if a.size else 0.0
    amax = float(a.max()) if a.size else 0.0
    struct.pack_into("<10i6f3i3f2i", hdr, 0, nx, ny, nz, 2, 0, 0, 0, nx, ny, nz,
                     nx * apix, ny * apix, nz * apix, 90.0, 90.0, 90.0, 1, 2, 3,
                     amin, amax, mean, 0 if nz == 1 or nz != nx else 1, 0)
    struct.pack_into("<3f", hdr, 196, 0.0, 0.0, 0.0)
    hdr[208:212] = b"MAP "
    hdr[212:216] = bytes([0x44, 0x44, 0x00, 0x00])
    struct.pack_into("<f", hdr, 216, rms)
    struct.pack_into("<i", hdr, 220, 1)
    lab = label.encode()[:80]
    hdr[224:224 + len(lab)] = lab
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(a.tobytes())
