"""C++ file codecs of the drop-in (MRC/DV and TIFF), checked on the CPU.

The C++ readers and writers (cudasirecon_amd/csrc/io_mrc.cpp, io_tiff.cpp)
are what the cudasirecon executable and SR_* ABI use in place of IVE and
CImg/libtiff (src/cudaSirecon/mrc.h:18-96, cudaSirecon.cpp:1118-1135,
1532-1617, 1636-1700).  They are compared with independent readers: Pillow's
libtiff-backed TIFF codec and the Python MRC codec, and with files whose
section order is written out by explicit nested loops.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from cases import GOLDEN
from cudasirecon_amd import mrc

PIL = pytest.importorskip("PIL.Image")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "cudasirecon_amd", "csrc")


@pytest.fixture(scope="module")
def codec(tmp_path_factory):
    d = tmp_path_factory.mktemp("codec")
    exe = str(d / "codec_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", CSRC, os.path.join(ROOT, "tests", "native", "codec_check.cpp"),
                    os.path.join(CSRC, "io_mrc.cpp"), os.path.join(CSRC, "io_tiff.cpp"), "-lz", "-o", exe],
                   check=True)

    def run(*args):
        r = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stderr.strip())
        return tuple(int(v) for v in r.stdout.split())

    run.dir = d
    return run


def pil_frames(path):
    im = PIL.open(path)
    out = []
    for i in range(getattr(im, "n_frames", 1)):
        im.seek(i)
        out.append(np.array(im))
    return np.stack(out)


def test_mrc_golden_sections_match_python_codec(codec):
    out = codec.dir / "raw.f32"
    nx, ny, nz = codec("mrc-read", os.path.join(GOLDEN, "raw.dv"), out, 0, 0)
    h, ref = mrc.read(os.path.join(GOLDEN, "raw.dv"))
    got = np.fromfile(out, np.float32).reshape(nz, ny, nx)
    assert (nx, ny, nz) == (h["nx"], h["ny"], h["nz"])
    np.testing.assert_array_equal(got, ref.astype(np.float32))


def test_mrc_complex_otf_matches_python_codec(codec):
    out = codec.dir / "otf.c64"
    nx, ny, nz = codec("mrc-c64", os.path.join(GOLDEN, "otf.dv"), out)
    h, ref = mrc.read(os.path.join(GOLDEN, "otf.dv"))
    got = np.fromfile(out, np.complex64).reshape(nz, ny, nx)
    np.testing.assert_array_equal(got, ref.astype(np.complex64))


def _mrc_bytes(nx, ny, sections, mode, order, nw=1, nt=1, interleaved=0):
    """A DV file written field by field in the given byte order."""
    hdr = bytearray(1024)
    e = order
    struct.pack_into(e + "3i", hdr, 0, nx, ny, len(sections))
    struct.pack_into(e + "i", hdr, 12, mode)
    struct.pack_into(e + "3f", hdr, 40, 0.08, 0.08, 0.125)
    struct.pack_into(e + "h", hdr, 96, -16224)
    struct.pack_into(e + "2h", hdr, 180, nt, interleaved)
    struct.pack_into(e + "2h", hdr, 196, nw, 515)
    dt = {1: "i2", 2: "f4", 6: "u2"}[mode]
    body = b"".join(np.asarray(s, dtype=np.dtype(e + dt)).tobytes() for s in sections)
    return bytes(hdr) + body


@pytest.mark.parametrize("order", ["<", ">"])
@pytest.mark.parametrize("mode", [1, 2, 6])
def test_mrc_byte_order_and_modes(codec, order, mode):
    rng = np.random.default_rng(3)
    secs = [rng.integers(0, 3000, size=(6, 5)) for _ in range(4)]
    p = codec.dir / ("m%d%s.dv" % (mode, "le" if order == "<" else "be"))
    p.write_bytes(_mrc_bytes(5, 6, secs, mode, order))
    out = codec.dir / "m.f32"
    assert codec("mrc-read", p, out, 0, 0) == (5, 6, 4)
    np.testing.assert_array_equal(np.fromfile(out, np.float32).reshape(4, 6, 5), np.array(secs, np.float32))


@pytest.mark.parametrize("interleaved", [0, 1, 2])
def test_mrc_wave_time_interleave(codec, interleaved):
    # sections written by explicit nested loops; the value encodes (z, w, t)
    nzz, nw, nt = 3, 2, 4
    loops = {0: ("w", "t", "z"), 1: ("t", "z", "w"), 2: ("t", "w", "z")}[interleaved]  # outer -> inner
    size = {"z": nzz, "w": nw, "t": nt}
    secs = []
    idx = {}
    for a in range(size[loops[0]]):
        for b in range(size[loops[1]]):
            for c in range(size[loops[2]]):
                idx.update({loops[0]: a, loops[1]: b, loops[2]: c})
                secs.append(np.full((2, 3), 100 * idx["z"] + 10 * idx["w"] + idx["t"]))
    p = codec.dir / ("il%d.dv" % interleaved)
    p.write_bytes(_mrc_bytes(3, 2, secs, 2, "<", nw=nw, nt=nt, interleaved=interleaved))
    out = codec.dir / "il.f32"
    for w in range(nw):
        for t in range(nt):
            assert codec("mrc-read", p, out, w, t) == (3, 2, nzz)
            got = np.fromfile(out, np.float32).reshape(nzz, 2, 3)[:, 0, 0]
            np.testing.assert_array_equal(got, [100 * z + 10 * w + t for z in range(nzz)])


def test_mrc_writer_read_by_python_codec(codec):
    rng = np.random.default_rng(4)
    vol = rng.standard_normal((3, 7, 9)).astype(np.float32)
    src = codec.dir / "w.f32"
    vol.tofile(src)
    dst = codec.dir / "w.dv"
    codec("mrc-write", src, 9, 7, 3, dst)
    h, got = mrc.read(str(dst))
    assert (h["nx"], h["ny"], h["nz"], h["mode"], h["dvid"]) == (9, 7, 3, 2, -16224)
    assert abs(h["xlen"] - 0.04) < 1e-7 and h["iwav1"] == 515
    np.testing.assert_array_equal(got, vol)


def test_tiff_golden_matches_pillow(codec):
    for name in ("raw.tif", "otf.tif"):
        out = codec.dir / (name + ".f32")
        nx, ny, nz = codec("tiff-read", os.path.join(GOLDEN, name), out)
        ref = pil_frames(os.path.join(GOLDEN, name)).astype(np.float32)
        assert ref.shape == (nz, ny, nx)
        np.testing.assert_array_equal(np.fromfile(out, np.float32).reshape(nz, ny, nx), ref)


@pytest.mark.parametrize("compression", [None, "tiff_lzw", "tiff_adobe_deflate", "packbits"])
def test_tiff_u16_compressions_match_pillow(codec, compression):
    rng = np.random.default_rng(5)
    # smooth plus noise so LZW/Deflate see both runs and literals
    base = (np.arange(40)[None, :] * 37 + np.arange(33)[:, None] * 11) % 4000
    frames = [(base + rng.integers(0, 50, size=base.shape) * (i % 2)).astype(np.uint16) for i in range(5)]
    p = codec.dir / ("c_%s.tif" % compression)
    ims = [PIL.fromarray(f) for f in frames]
    kw = {"compression": compression} if compression else {}
    ims[0].save(p, save_all=True, append_images=ims[1:], **kw)
    out = codec.dir / "c.u16"
    assert codec("tiff-u16", p, out) == (40, 33, 5)
    got = np.fromfile(out, np.uint16).reshape(5, 33, 40)
    np.testing.assert_array_equal(got, pil_frames(p))
    np.testing.assert_array_equal(got, np.array(frames))


def test_tiff_float_lzw_matches_pillow(codec):
    rng = np.random.default_rng(6)
    frames = [rng.standard_normal((17, 23)).astype(np.float32) for _ in range(3)]
    p = codec.dir / "f_lzw.tif"
    ims = [PIL.fromarray(f, mode="F") for f in frames]
    ims[0].save(p, save_all=True, append_images=ims[1:], compression="tiff_lzw")
    out = codec.dir / "f.f32"
    assert codec("tiff-read", p, out) == (23, 17, 3)
    np.testing.assert_array_equal(np.fromfile(out, np.float32).reshape(3, 17, 23), np.array(frames))


def test_tiff_big_endian_strips(codec):
    # hand-made big-endian uint16 stack, 2 strips per page
    nx, ny, nz = 4, 6, 2
    data = (np.arange(nz * ny * nx) * 7).astype(">u2").reshape(nz, ny, nx)
    ents = 9
    ifd_size = 2 + ents * 12 + 4
    buf = bytearray(b"MM\x00\x2a" + struct.pack(">I", 8))
    pix_start = 8 + nz * (ifd_size + 16)
    for z in range(nz):
        ifd = 8 + z * (ifd_size + 16)
        arr_off = ifd + ifd_size  # strip offsets (2 longs) then byte counts (2 longs)
        page_off = pix_start + z * nx * ny * 2
        e = [(256, 3, 1, nx), (257, 3, 1, ny), (258, 3, 1, 16), (259, 3, 1, 1), (262, 3, 1, 1),
             (273, 4, 2, arr_off), (277, 3, 1, 1), (278, 3, 1, 3), (279, 4, 2, arr_off + 8)]
        rec = struct.pack(">H", ents)
        for tag, typ, cnt, val in e:
            if typ == 3 and cnt == 1:
                rec += struct.pack(">HHIHH", tag, typ, cnt, val, 0)
            else:
                rec += struct.pack(">HHII", tag, typ, cnt, val)
        nxt = 8 + (z + 1) * (ifd_size + 16) if z + 1 < nz else 0
        rec += struct.pack(">I", nxt)
        half = nx * 3 * 2
        rec += struct.pack(">IIII", page_off, page_off + half, half, half)
        buf += rec
    buf += data.tobytes()
    p = codec.dir / "be.tif"
    p.write_bytes(bytes(buf))
    out = codec.dir / "be.u16"
    assert codec("tiff-u16", p, out) == (nx, ny, nz)
    np.testing.assert_array_equal(np.fromfile(out, np.uint16).reshape(nz, ny, nx), data.astype(np.uint16))


def test_tiff_writer_read_by_pillow(codec):
    rng = np.random.default_rng(7)
    vol = rng.standard_normal((4, 9, 11)).astype(np.float32)
    src = codec.dir / "tw.f32"
    vol.tofile(src)
    dst = codec.dir / "tw.tif"
    codec("tiff-write", src, 11, 9, 4, dst)
    np.testing.assert_array_equal(pil_frames(dst), vol)
    # and by the C++ reader
    out = codec.dir / "tw_back.f32"
    assert codec("tiff-read", dst, out) == (11, 9, 4)
    np.testing.assert_array_equal(np.fromfile(out, np.float32).reshape(4, 9, 11), vol)


def test_tiff_rejects_truncated(codec):
    p = codec.dir / "trunc.tif"
    p.write_bytes(open(os.path.join(GOLDEN, "otf.tif"), "rb").read()[:5000])
    with pytest.raises(RuntimeError, match="TIFF"):
        codec("tiff-read", p, codec.dir / "x.f32")


def _bigtiff_u16(path, strip_off, strip_cnt, width=4, height=4, extra=()):
    """A one-page BigTIFF (little endian) whose strip offset and byte count are
    given as 64-bit values, plus 16 bytes of pixel data after the IFD."""
    entries = [(256, 3, 1, width), (257, 3, 1, height), (258, 3, 1, 16), (259, 3, 1, 1), (277, 3, 1, 1),
               (278, 3, 1, height), (273, 16, 1, strip_off), (279, 16, 1, strip_cnt)] + list(extra)
    entries.sort()
    ifd = struct.pack("<Q", len(entries))
    for tag, typ, cnt, val in entries:
        ifd += struct.pack("<HHQQ", tag, typ, cnt, val)
    ifd += struct.pack("<Q", 0)
    data = struct.pack("<2sHHHQ", b"II", 43, 8, 0, 16) + ifd + bytes(range(32))
    path.write_bytes(data)


def test_tiff_offsets_near_2_64_are_refused(codec):
    """Offsets and counts chosen so that offset + count wraps past 2^64 to a
    small number: the reader must refuse them, not read out of bounds."""
    good = codec.dir / "big_ok.tif"
    n_hdr = 16 + 8 + 8 * 20 + 8
    _bigtiff_u16(good, n_hdr, 32)
    assert codec("tiff-u16", good, codec.dir / "ok.u16") == (4, 4, 1)
    bad = codec.dir / "big_wrap.tif"
    _bigtiff_u16(bad, 2 ** 64 - 16, 32)  # offset + count wraps to 16
    with pytest.raises(RuntimeError, match="past the end"):
        codec("tiff-u16", bad, codec.dir / "x.u16")
    bad2 = codec.dir / "big_count.tif"  # an extra entry whose count * 8 wraps to 0 bytes
    _bigtiff_u16(bad2, n_hdr + 20, 32, extra=[(324, 16, 2 ** 61, 0)])
    with pytest.raises(RuntimeError, match="past the end"):
        codec("tiff-u16", bad2, codec.dir / "y.u16")
