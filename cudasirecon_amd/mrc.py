"""MRC / DeltaVision (Priism "IVE") file codec.

Reads and writes the 1024-byte Priism header fields the reference uses
(src/cudaSirecon/mrc.h:18-60; cudaSirecon.cpp:1485-1522): nx, ny, nz, mode,
xlen/ylen/zlen, amin/amax/amean, the extended-header size, num_times,
interleave, num_waves, iwav1-5 and the 10 label lines.  Both byte orders are
accepted (the DV stamp -16224 at offset 96 identifies the order).
The C++ codec used by the cudasirecon executable (cudasirecon_amd/csrc/io_mrc.cpp)
implements the same layout; tests/test_io.py checks that both agree.
"""
import struct

import numpy as np

MODES = {0: np.uint8, 1: np.int16, 2: np.float32, 3: np.complex64, 4: np.complex64, 6: np.uint16}
DV_STAMP = -16224

# (name, struct code, offset)
_FIELDS = [
    ("nx", "i", 0), ("ny", "i", 4), ("nz", "i", 8), ("mode", "i", 12),
    ("nxst", "i", 16), ("nyst", "i", 20), ("nzst", "i", 24),
    ("mx", "i", 28), ("my", "i", 32), ("mz", "i", 36),
    ("xlen", "f", 40), ("ylen", "f", 44), ("zlen", "f", 48),
    ("alpha", "f", 52), ("beta", "f", 56), ("gamma", "f", 60),
    ("mapc", "i", 64), ("mapr", "i", 68), ("maps", "i", 72),
    ("amin", "f", 76), ("amax", "f", 80), ("amean", "f", 84),
    ("ispg", "i", 88), ("next", "i", 92), ("dvid", "h", 96), ("nblank", "h", 98),
    ("ntst", "i", 100), ("nint", "h", 128), ("nreal", "h", 130),
    ("nres", "h", 132), ("nzfact", "h", 134),
    ("file_type", "h", 160), ("lens", "h", 162),
    ("num_times", "h", 180), ("interleaved", "h", 182),
    ("num_waves", "h", 196), ("iwav1", "h", 198), ("iwav2", "h", 200), ("iwav3", "h", 202),
    ("iwav4", "h", 204), ("iwav5", "h", 206),
    ("zorig", "f", 208), ("xorig", "f", 212), ("yorig", "f", 216), ("nlab", "i", 220),
]


def _endian(raw):
    le = struct.unpack_from("<h", raw, 96)[0]
    if le == DV_STAMP:
        return "<"
    be = struct.unpack_from(">h", raw, 96)[0]
    if be == DV_STAMP:
        return ">"
    # plain MRC without the DV stamp: guess from a sane mode value
    mode_le = struct.unpack_from("<i", raw, 12)[0]
    return "<" if 0 <= mode_le <= 16 else ">"


def read_header(raw):
    if len(raw) < 1024:
        raise ValueError("MRC: file shorter than the 1024-byte header")
    e = _endian(raw)
    h = {"byteorder": e}
    for name, code, off in _FIELDS:
        h[name] = struct.unpack_from(e + code, raw, off)[0]
    h["labels"] = [raw[224 + 80 * i: 224 + 80 * (i + 1)].decode("latin-1") for i in range(10)]
    if h["num_waves"] <= 0:
        h["num_waves"] = 1
    if h["num_times"] <= 0:
        h["num_times"] = 1
    return h


def section_index(h, z, w, t, nz_per=None):
    """Section number of (z, wave, time) for the header's interleave (IMPosnZWT).

    interleaved: 0 = Z fastest then T then W (ZTW), 1 = W fastest then Z then T (WZT),
    2 = Z fastest then W then T (ZWT).
    """
    nw, nt = h["num_waves"], h["num_times"]
    nzz = nz_per if nz_per is not None else h["nz"] // (nw * nt)
    il = h["interleaved"]
    if il == 1:
        return (t * nzz + z) * nw + w
    if il == 2:
        return (t * nw + w) * nzz + z
    return (w * nt + t) * nzz + z


def read(path):
    """Return (header dict, data array [nz][ny][nx] in the file's numeric type)."""
    with open(path, "rb") as f:
        raw = f.read()
    h = read_header(raw)
    dt = MODES.get(h["mode"])
    if dt is None:
        raise ValueError("MRC: unsupported mode %d" % h["mode"])
    dt = np.dtype(dt).newbyteorder(h["byteorder"])
    start = 1024 + h["next"]
    n = h["nx"] * h["ny"] * h["nz"]
    if len(raw) < start + n * dt.itemsize:
        raise ValueError("MRC: file truncated")
    data = np.frombuffer(raw, dtype=dt, count=n, offset=start).reshape(h["nz"], h["ny"], h["nx"])
    return h, data.astype(dt.newbyteorder("="))


def make_header(nx, ny, nz, mode, xlen=0.1, ylen=0.1, zlen=0.2, num_waves=1, num_times=1,
                wave=530, interleaved=0, labels=(), amin=0.0, amax=0.0, amean=0.0):
    h = {name: 0 for name, _, _ in _FIELDS}
    h.update(nx=nx, ny=ny, nz=nz, mode=mode, mx=nx, my=ny, mz=nz, xlen=xlen, ylen=ylen, zlen=zlen,
             alpha=90.0, beta=90.0, gamma=90.0, mapc=1, mapr=2, maps=3, dvid=DV_STAMP,
             num_waves=num_waves, num_times=num_times, iwav1=wave, interleaved=interleaved,
             amin=amin, amax=amax, amean=amean, nlab=len(labels))
    h["labels"] = list(labels)
    h["byteorder"] = "<"
    return h


def write(path, data, header):
    data = np.asarray(data)
    mode = header["mode"]
    dt = np.dtype(MODES[mode]).newbyteorder("<")
    buf = bytearray(1024)
    for name, code, off in _FIELDS:
        v = header.get(name, 0)
        if name == "next":
            v = 0
        struct.pack_into("<" + code, buf, off, v)
    for i, lab in enumerate(header.get("labels", [])[:10]):
        b = lab.encode("latin-1")[:80].ljust(80, b" ")
        buf[224 + 80 * i: 224 + 80 * (i + 1)] = b
    with open(path, "wb") as f:
        f.write(bytes(buf))
        f.write(np.ascontiguousarray(data, dtype=dt).tobytes())
