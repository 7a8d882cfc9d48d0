"""MRC2014 reading and writing (the layout of libEM/io/mrcio.cpp): header
words at their byte offsets, extended headers skipped, both byte orders,
integer modes, stacks memory-mapped."""
import struct

import numpy as np

from eman2_amd import mrc


def test_round_trip_and_header_offsets(tmp_path):
    p = tmp_path / "map.mrc"
    a = np.random.default_rng(1).standard_normal((6, 5, 4)).astype(np.float32)
    mrc.write(str(p), a, apix=1.5)
    raw = p.read_bytes()
    assert len(raw) == 1024 + a.nbytes
    assert struct.unpack("<4i", raw[0:16]) == (4, 5, 6, 2)          # nx ny nz mode
    assert struct.unpack("<3i", raw[28:40]) == (4, 5, 6)            # mx my mz
    assert struct.unpack("<3f", raw[40:52]) == (6.0, 7.5, 9.0)      # cell = apix * sampling
    assert struct.unpack("<3i", raw[64:76]) == (1, 2, 3)            # x fastest
    assert raw[208:212] == b"MAP " and raw[212:214] == b"\x44\x44"
    b, h = mrc.read(str(p))
    assert h["apix"] == 1.5 and b.shape == (6, 5, 4)
    np.testing.assert_array_equal(np.asarray(b), a)
    assert abs(h["dmean"] - a.mean()) < 1e-6


def _hand_made(path, data, mode, end="<", nsymbt=0, apix=2.0):
    nz, ny, nx = data.shape
    hdr = bytearray(1024)
    struct.pack_into(end + "10i6f3i3f2i", hdr, 0, nx, ny, nz, mode, 0, 0, 0, nx, ny, nz,
                     nx * apix, ny * apix, nz * apix, 90, 90, 90, 1, 2, 3, 0, 0, 0, 0, nsymbt)
    hdr[208:212] = b"MAP "
    hdr[212:214] = b"\x11\x11" if end == ">" else b"\x44\x44"
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(b"\x7f" * nsymbt)
        f.write(data.astype(np.dtype(data.dtype).newbyteorder(end)).tobytes())


def test_extended_header_big_endian_and_int16(tmp_path):
    d = (np.arange(2 * 3 * 4, dtype=np.int16) - 7).reshape(2, 3, 4)
    p = tmp_path / "be.mrcs"
    _hand_made(str(p), d, 1, end=">", nsymbt=96)
    b, h = mrc.read(str(p), mmap=False)
    assert h["nsymbt"] == 96 and h["apix"] == 2.0
    np.testing.assert_array_equal(b, d)


def test_stack_is_memory_mapped(tmp_path):
    p = tmp_path / "stack.mrcs"
    a = np.random.default_rng(2).random((10, 8, 8)).astype(np.float32)
    mrc.write(str(p), a, apix=1.0)
    b, h = mrc.read(str(p))
    assert isinstance(b, np.memmap) and h["nz"] == 10
    np.testing.assert_array_equal(np.asarray(b[3]), a[3])
