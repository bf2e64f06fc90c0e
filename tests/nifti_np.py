"""A small NIfTI-1 reader/writer in numpy for the tests: an implementation of
the public format independent of the engine's C++ one (fdt_amd/csrc/fdt_nifti.cpp)."""
import gzip
import struct

import numpy as np

DT = {2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 64: np.float64, 256: np.int8, 512: np.uint16,
      768: np.uint32}
CODE = {np.dtype(v): k for k, v in DT.items()}


def write(path, arr, pixdim=(2.0, 2.0, 2.0), slope=0.0, inter=0.0, big_endian=False, sform=None, cal=(0.0, 0.0)):
    """arr indexed [x, y, z(, t)]"""
    arr = np.asarray(arr)
    e = ">" if big_endian else "<"
    h = bytearray(352)
    struct.pack_into(e + "i", h, 0, 348)
    dims = [arr.ndim] + list(arr.shape) + [1] * (7 - arr.ndim)
    struct.pack_into(e + "8h", h, 40, *dims)
    struct.pack_into(e + "h", h, 70, CODE[arr.dtype])
    struct.pack_into(e + "h", h, 72, arr.dtype.itemsize * 8)
    pd = [1.0] + list(pixdim) + [1.0] * (7 - len(pixdim))
    struct.pack_into(e + "8f", h, 76, *pd)
    struct.pack_into(e + "f", h, 108, 352.0)
    struct.pack_into(e + "f", h, 112, slope)
    struct.pack_into(e + "f", h, 116, inter)
    struct.pack_into(e + "f", h, 124, cal[1])
    struct.pack_into(e + "f", h, 128, cal[0])
    if sform is not None:
        struct.pack_into(e + "h", h, 254, 1)
        struct.pack_into(e + "12f", h, 280, *np.asarray(sform, dtype=np.float32).ravel())
    h[344:348] = b"n+1\0"
    body = np.asfortranarray(arr).astype(arr.dtype.newbyteorder(e)).tobytes(order="F")
    raw = bytes(h) + body
    if str(path).endswith(".gz"):
        with gzip.open(path, "wb") as f:
            f.write(raw)
    else:
        with open(path, "wb") as f:
            f.write(raw)


def read(path):
    """-> (float32 array [x, y, z(, t)], header dict)"""
    op = gzip.open if str(path).endswith(".gz") else open
    with op(path, "rb") as f:
        raw = f.read()
    e = "<" if struct.unpack_from("<i", raw, 0)[0] == 348 else ">"
    dims = struct.unpack_from(e + "8h", raw, 40)
    dt = struct.unpack_from(e + "h", raw, 70)[0]
    pixdim = struct.unpack_from(e + "8f", raw, 76)
    off = int(struct.unpack_from(e + "f", raw, 108)[0])
    slope, inter = struct.unpack_from(e + "2f", raw, 112)
    cal_max, cal_min = struct.unpack_from(e + "2f", raw, 124)
    sform_code = struct.unpack_from(e + "h", raw, 254)[0]
    srow = struct.unpack_from(e + "12f", raw, 280)
    shape = dims[1:1 + dims[0]]
    n = int(np.prod(shape))
    a = np.frombuffer(raw, dtype=np.dtype(DT[dt]).newbyteorder(e), count=n, offset=off).reshape(shape, order="F")
    a = a.astype(np.float32)
    if slope != 0 and not (slope == 1 and inter == 0):
        a = a * np.float32(slope) + np.float32(inter)
    return a, dict(dims=dims, datatype=dt, pixdim=pixdim, cal=(cal_min, cal_max), sform_code=sform_code, srow=srow,
                   magic=raw[344:348])
