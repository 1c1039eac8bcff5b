"""Minimal single-file NIfTI-1 writer / reader for the command line tests,
written from the NIfTI-1 header layout (348 bytes, magic "n+1", voxel data
at vox_offset, x fastest)."""
import gzip
import struct

import numpy as np

CODES = {np.dtype(np.uint8): 2, np.dtype(np.int16): 4, np.dtype(np.int32): 8, np.dtype(np.float32): 16,
         np.dtype(np.float64): 64, np.dtype(np.uint16): 512}


def _header(shape_xyzt, code, bitpix, order):
    dims = [len(shape_xyzt)] + list(shape_xyzt) + [1] * (7 - len(shape_xyzt))
    h = bytearray(348)
    struct.pack_into(order + "i", h, 0, 348)
    struct.pack_into(order + "8h", h, 40, *dims)
    struct.pack_into(order + "hh", h, 70, code, bitpix)
    struct.pack_into(order + "8f", h, 76, 1.0, 2.0, 2.0, 2.0, 1.0, 1.0, 1.0, 1.0)
    struct.pack_into(order + "f", h, 108, 352.0)       # vox_offset
    struct.pack_into(order + "ff", h, 112, 1.0, 0.0)   # scl_slope, scl_inter
    h[344:348] = b"n+1\0"
    return bytes(h)


def write(path, vol, big_endian=False):
    """vol: numpy array indexed [t][z][y][x] (or [z][y][x])."""
    vol = np.asarray(vol)
    order = ">" if big_endian else "<"
    shape = vol.shape[::-1]
    hdr = _header(shape, CODES[vol.dtype], vol.dtype.itemsize * 8, order)
    data = vol.astype(vol.dtype.newbyteorder(order)).tobytes()
    blob = hdr + b"\0\0\0\0" + data
    if str(path).endswith(".gz"):
        with gzip.open(path, "wb") as f:
            f.write(blob)
    else:
        with open(path, "wb") as f:
            f.write(blob)


def read(path):
    """Returns (array [t][z][y][x] float32 view of the stored type, header dict)."""
    raw = gzip.open(path).read() if str(path).endswith(".gz") else open(path, "rb").read()
    order = "<" if struct.unpack_from("<i", raw, 0)[0] == 348 else ">"
    dims = struct.unpack_from(order + "8h", raw, 40)
    code, bitpix = struct.unpack_from(order + "hh", raw, 70)
    off = int(struct.unpack_from(order + "f", raw, 108)[0])
    cal_max, cal_min = struct.unpack_from(order + "ff", raw, 124)
    dt = {v: k for k, v in CODES.items()}[code].newbyteorder(order)
    shape = [d for d in dims[1:dims[0] + 1]]
    n = int(np.prod(shape))
    arr = np.frombuffer(raw, dt, count=n, offset=off).reshape(shape[::-1])
    return arr, dict(dims=dims, datatype=code, bitpix=bitpix, cal_min=cal_min, cal_max=cal_max,
                     magic=raw[344:348])
