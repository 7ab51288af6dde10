"""GPU parity of the drop-in boundary: the cudasirecon executable and the SR_* ABI.

Runs the reference's own test_data invocations (test_data/run.sh: MRC mode
with config, TIFF-folder mode with config-tiff) through the executable, and
the in-memory SR_new_from_shape path through the ABI, and compares every
result with the FP32 CPU oracle on the same inputs (output rel-L2 <= 1e-5).
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from cases import GOLDEN, full_otf_from_ra, load_testdata, load_testdata_tiff, rel_l2, tiff_frames
from cudasirecon_amd import mrc, sr
from oracle import binding as ob

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cudasirecon_amd", "bin", "cudasirecon")
TOL = 1e-5


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_exe(args, timeout=120):
    log("run", " ".join(os.path.basename(a) for a in args))
    r = subprocess.run([EXE] + args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr
    assert "!!Error occurred" not in r.stderr, r.stderr
    return r


def oracle_series(p, otf, raws, fit64=True):
    o = ob.Oracle(p, otf, 32, fit64)
    out = []
    for t, raw in enumerate(raws):
        o.load(raw)
        o.rescale(t)
        o.set_time(t)
        o.process_volume()
        out.append(o.result())
    return out


@pytest.fixture(scope="module")
def oracle_mrc():
    p, otf, raw = load_testdata()
    return oracle_series(p, otf, [raw])[0]


@pytest.fixture(scope="module")
def oracle_tiff():
    p, otf, raw = load_testdata_tiff()
    return oracle_series(p, otf, [raw])[0]


def test_cli_mrc_matches_oracle(tmp_path, oracle_mrc):
    out = str(tmp_path / "proc.dv")
    run_exe([os.path.join(GOLDEN, "raw.dv"), out, os.path.join(GOLDEN, "otf.dv"), "-c", os.path.join(GOLDEN, "config")])
    h, vol = mrc.read(out)
    hin = mrc.read_header(open(os.path.join(GOLDEN, "raw.dv"), "rb").read(1024))
    # setOutputHeader (src/cudaSirecon/mrc.h:47-60)
    assert (h["nx"], h["ny"], h["nz"], h["mode"], h["next"]) == (128, 128, 9, 2, 0)
    assert h["xlen"] == pytest.approx(hin["xlen"] / 2) and h["ylen"] == pytest.approx(hin["ylen"] / 2)
    assert h["zlen"] == pytest.approx(hin["zlen"]) and h["iwav1"] == hin["iwav1"]
    err = rel_l2(vol, oracle_mrc)
    log("cli mrc rel-L2 %.3e" % err)
    assert err < TOL
    # amin/amax of time point 0 with the reference's else-if scan (cudaSirecon.cpp:1671-1689)
    mx, mn = -np.float32(np.finfo(np.float32).max), np.float32(np.finfo(np.float32).max)
    flat = vol.reshape(-1)
    for v in flat:
        if v > mx:
            mx = v
        elif v < mn:
            mn = v
    assert h["amax"] == mx and h["amin"] == mn
    assert h["amean"] == hin["amean"]


def test_cli_tiff_folder_matches_oracle(tmp_path, oracle_tiff):
    d = tmp_path / "data"
    d.mkdir()
    shutil.copy(os.path.join(GOLDEN, "raw.tif"), d / "raw.tif")
    run_exe([str(d), "raw", os.path.join(GOLDEN, "otf.tif"), "-c", os.path.join(GOLDEN, "config-tiff")])
    out = d / "GPUsirecon" / "raw_proc.tif"
    assert out.exists()
    vol = tiff_frames(str(out))
    assert vol.shape == (9, 128, 128) and vol.dtype == np.float32
    err = rel_l2(vol, oracle_tiff)
    log("cli tiff rel-L2 %.3e" % err)
    assert err < TOL


def _shape_config(tmp_path, extra=""):
    cfg = tmp_path / "config-mem"
    cfg.write_text(open(os.path.join(GOLDEN, "config-tiff")).read() + "\notf-file=%s\n%s"
                   % (os.path.join(GOLDEN, "otf.tif"), extra))
    return str(cfg)


def test_sr_in_memory_matches_oracle_and_cli(tmp_path, oracle_tiff):
    p, otf, raw = load_testdata_tiff()
    with sr.SIMReconstructor(shape=(64, 64, raw.shape[0]), config=_shape_config(tmp_path)) as r:
        im = r.image_params
        assert (im.nx, im.ny, im.nz, im.ntimes) == (64, 64, 9, 1)
        assert r.recon_params.bTIFF == 1 and r.recon_params.nxotf == 129 and r.recon_params.nzotf == 65
        vol = r.reconstruct(raw)
    err = rel_l2(vol, oracle_tiff)
    log("SR in-memory rel-L2 %.3e" % err)
    assert err < TOL
    # the CLI's TIFF mode reads the same float pages: identical result
    d = tmp_path / "data"
    d.mkdir()
    shutil.copy(os.path.join(GOLDEN, "raw.tif"), d / "raw.tif")
    run_exe([str(d), "raw", os.path.join(GOLDEN, "otf.tif"), "-c", os.path.join(GOLDEN, "config-tiff")])
    np.testing.assert_array_equal(tiff_frames(str(d / "GPUsirecon" / "raw_proc.tif")), vol)


def test_sr_parameter_change_takes_effect(tmp_path):
    p, otf, raw = load_testdata_tiff()
    with sr.SIMReconstructor(shape=(64, 64, raw.shape[0]), config=_shape_config(tmp_path)) as r:
        first = r.reconstruct(raw)
        r.recon_params.wiener = 0.01
        r.recon_params.bDampenOrder0 = 0
        second = r.reconstruct(raw)
    q = p.copy()
    q.wiener = 0.01
    q.bDampenOrder0 = 0
    ref = oracle_series(q, otf, [raw])[0]
    err = rel_l2(second, ref)
    log("SR after parameter change rel-L2 %.3e" % err)
    assert err < TOL and rel_l2(first, second) > 1e-3


def test_two_handles_in_one_process(tmp_path):
    _, _, raw = load_testdata_tiff()
    cfg = _shape_config(tmp_path)
    a = sr.SIMReconstructor(shape=(64, 64, raw.shape[0]), config=cfg)
    b = sr.SIMReconstructor(shape=(64, 64, raw.shape[0]), config=cfg)
    a.set_raw(raw)
    b.set_raw(raw * np.float32(0.5))
    for h in (a, b):
        h.load_and_rescale_image(0, 0)
        h.set_cur_time_idx(0)
        h.process_one_volume()
    va, vb = a.get_result(), b.get_result()
    a.close()
    b.close()
    np.testing.assert_array_equal(va, sr.SIMReconstructor(shape=(64, 64, raw.shape[0]), config=cfg).reconstruct(raw))
    # input scale cancels in the Wiener-filtered, modamp-normalised result only up to
    # the bleach/rescale path, so just check it is a distinct, finite volume
    assert np.isfinite(vb).all() and rel_l2(vb, va) > 0


def test_cli_mrc_time_series_matches_oracle(tmp_path):
    """Two time points in one DV file: t>0 reuses the time-0 k0 and refits modamps."""
    p, otf, raw = load_testdata()
    h, data = mrc.read(os.path.join(GOLDEN, "raw.dv"))
    rng = np.random.default_rng(9)
    t1 = np.clip(data.astype(np.float32) * 0.9 + rng.normal(0, 2, data.shape), 0, 65535).astype(data.dtype)
    hh = dict(h)
    hh["nz"] = 2 * h["nz"]
    hh["num_times"] = 2
    hh["interleaved"] = 0
    src = str(tmp_path / "series.dv")
    mrc.write(src, np.concatenate([data, t1]), hh)
    out = str(tmp_path / "series_proc.dv")
    run_exe([src, out, os.path.join(GOLDEN, "otf.dv"), "-c", os.path.join(GOLDEN, "config")])
    ho, vol = mrc.read(out)
    assert ho["nz"] == 18 and vol.shape == (18, 128, 128)
    q = p.copy()
    q.ntimes = 2
    refs = oracle_series(q, otf, [raw, t1.astype(np.float32)])
    for t in range(2):
        err = rel_l2(vol[9 * t: 9 * (t + 1)], refs[t])
        log("time %d rel-L2 %.3e" % (t, err))
        assert err < TOL


def _resampled_case(extra_cfg):
    """(oracle params, otf, raw) for the test_data TIFF with deskew/crop options applied
    the way setup_common and loadImageData derive them (cudaSirecon.cpp:1153-1171, 1355-1360)."""
    import math
    p, otf, raw = load_testdata_tiff()
    q = p.copy()
    q.nx_raw = p.nx
    if "deskew" in extra_cfg:
        ang = float(np.float32(extra_cfg["deskew"]))
        if ang < 0:
            ang = float(np.float32(ang + 180.0))  # setup_common folds negative angles
        a = ang * math.pi / 180.0
        dz_raw = float(np.float32(p.dz))
        nx = int(64 + p.nz * abs(math.cos(a)) * dz_raw / float(np.float32(p.dxy)))
        assert nx == 64  # a good dimension already: findOptimalDimension keeps it
        q.nx = 64
        q.dz = float(np.float32(dz_raw * abs(math.sin(a))))
        q.deskew_factor = float(np.float32(math.cos(a) * dz_raw / float(np.float32(p.dxy))))
    if "cropXY" in extra_cfg:
        q.nx = int(extra_cfg["cropXY"])
    return q, otf, raw


@pytest.mark.parametrize("opts", [{"deskew": 87.0}, {"cropXY": 48}, {"deskew": 87.0, "deskewshift": 3},
                                  {"deskew": -93.0}])
def test_sr_deskew_and_crop_match_oracle(tmp_path, opts):
    log("case: resampling", opts)
    q, otf, raw = _resampled_case(opts)
    if "deskewshift" in opts:
        q.deskew_shift = int(opts["deskewshift"])
    extra = "\n".join("%s=%s" % kv for kv in opts.items())
    with sr.SIMReconstructor(shape=(64, 64, raw.shape[0]), config=_shape_config(tmp_path, extra)) as r:
        im = r.image_params
        assert (im.nx, im.nx_raw) == (q.nx, 64)
        if "deskew" in opts:
            assert im.dz == pytest.approx(q.dz, rel=1e-6) and im.dz_raw == pytest.approx(0.125)
        vol = r.reconstruct(raw)
    ref = oracle_series(q, otf, [raw])[0]
    assert vol.shape == ref.shape == (9, 2 * q.ny, 2 * q.nx)
    err = rel_l2(vol, ref)
    log("resampled %s rel-L2 %.3e" % (opts, err))
    assert err < TOL


def test_sr_bessel_matches_oracle(tmp_path):
    """--bessel through the config file: the reference's Bessel-sheet branches
    (order-1 fit, overlap cutoff 0.01, order-0 weight 4, Bessel axial cutoffs,
    anisotropic apodization) against the oracle's."""
    log("case: bessel besselNA=0.65 besselExWave=0.488")
    p, otf, raw = load_testdata_tiff()
    q = p.copy()
    q.bBessel, q.BesselNA, q.BesselLambdaEx = 1, float(np.float32(0.65)), float(np.float32(0.488))
    with sr.SIMReconstructor(shape=(64, 64, raw.shape[0]),
                             config=_shape_config(tmp_path, "bessel=true\nbesselNA=0.65\nbesselExWave=0.488")) as r:
        assert r.recon_params.bBessel == 1
        vol = r.reconstruct(raw)
    ref = oracle_series(q, otf, [raw])[0]
    err = rel_l2(vol, ref)
    log("bessel rel-L2 %.3e" % err)
    assert err < TOL


def test_multichannel_driver_matches_oracle(tmp_path, oracle_mrc):
    """The multichannel driver (recon.py counterpart): a two-wave DV stored ZWT,
    each channel reconstructed by the executable with its own otf<wave>.otf and
    config<wave>.txt (the second with wiener 0.002), merged W fastest; every
    channel of the merged file matches the oracle of that channel."""
    from cudasirecon_amd import recon
    from cudasirecon_amd._abi import set_params

    p, otf, raw = load_testdata()
    h, data = mrc.read(os.path.join(GOLDEN, "raw.dv"))
    rng = np.random.default_rng(12)
    ch2 = np.clip(data.astype(np.float32) * 0.7 + rng.normal(0, 2, data.shape), 0, None).astype(data.dtype)
    hh = dict(h, nz=2 * h["nz"], num_waves=2, iwav2=608, interleaved=2)
    src = str(tmp_path / "cells.dv")
    mrc.write(src, np.concatenate([data, ch2]), hh)
    cfg = open(os.path.join(GOLDEN, "config")).read()
    assert "wiener=0.001" in cfg
    for w, text in ((528, cfg), (608, cfg.replace("wiener=0.001", "wiener=0.002"))):
        shutil.copy(os.path.join(GOLDEN, "otf.dv"), tmp_path / ("otf%d.otf" % w))
        (tmp_path / ("config%d.txt" % w)).write_text(text)
    log("multichannel driver on a 2-wave 64x64x9 DV")
    out = recon.reconstruct(src, otf_dir=str(tmp_path), config_dir=str(tmp_path), app=EXE)
    ho, vol = mrc.read(out)
    assert (ho["num_waves"], ho["interleaved"], ho["iwav1"], ho["iwav2"]) == (2, 1, 528, 608)
    chans = [vol[[mrc.section_index(ho, z, c, 0) for z in range(9)]] for c in range(2)]
    q = p.copy()
    set_params(q, wave=608, wiener=0.002)
    ref2 = oracle_series(q, otf, [ch2.astype(np.float32)])[0]
    for c, ref in enumerate((oracle_mrc, ref2)):
        err = rel_l2(chans[c], ref)
        log("channel %d rel-L2 %.3e" % (c, err))
        assert err < TOL
    assert rel_l2(chans[0], chans[1]) > 1e-2  # the two channels did differ


def _preprocessed_oracle(p, otf, raw):
    o = ob.Oracle(p, otf, 32, False)
    o.load(raw)
    o.rescale(0)
    o.set_time(0)
    o.preprocess_bands()
    return o


def test_cli_debug_dumps_match_oracle(tmp_path):
    """--saveprefiltered: every direction's separated 3D spectra, no k0 search;
    --saveoverlaps: findk0's overlap pair (order 0 / order 2 at the k0 guess) of
    every direction in real space; --savealignedraw: the header only.  Each
    stops before the filter and writes no result (cudaSirecon.cpp:269-287,
    330-352, 1179-1209, 1286-1288; src/cudaSirecon/mrc.h:76-96)."""
    p, otf, raw = load_testdata()
    o = _preprocessed_oracle(p, otf, raw)
    args = [os.path.join(GOLDEN, "raw.dv"), None, os.path.join(GOLDEN, "otf.dv"), "-c", os.path.join(GOLDEN, "config")]

    def run(tag, opt, path):
        out = str(tmp_path / ("%s_out.dv" % tag))
        r = run_exe(args[:1] + [out] + args[2:] + [opt, path])
        assert "Quit processing because either bSaveSeparated, bSaveAlignedRaw, or bSaveOverlaps is TRUE" in r.stdout
        assert os.path.getsize(out) == 1024  # the output file holds its header only
        return r

    sep = str(tmp_path / "sep.dv")
    r = run("sep", "--saveprefiltered", sep)
    assert "direction 0: k0" not in r.stdout  # no k0 search
    h, d = mrc.read(sep)
    assert (h["nx"], h["ny"], h["nz"], h["mode"]) == (33, 64, 135, 4)
    assert h["labels"][h["nlab"] - 1].startswith("separated bands of all directions")
    assert (h["amin"], h["amax"], h["amean"]) == (0.0, 1.0, 0.0)
    got = d.reshape(3, 5, 9, 64, 33)
    for dd in range(3):
        for b in range(5):
            err = rel_l2(got[dd, b], o.band(dd, b))
            assert err < TOL, (dd, b, err)
    log("saveprefiltered ok")

    ov = str(tmp_path / "ov.dv")
    r = run("ov", "--saveoverlaps", ov)
    assert "direction 2: k0" in r.stdout  # the fit ran
    h, d = mrc.read(ov)
    assert (h["nx"], h["ny"], h["nz"], h["mode"], h["num_waves"], h["interleaved"]) == (64, 64, 54, 4, 2, 1)
    assert h["labels"][h["nlab"] - 1].startswith("overlaps in real space")
    got = d.reshape(3, 9, 2, 64, 64)
    for dd in range(3):
        f = o.fit(dd)
        ref = o.makeoverlaps(dd, 0, 2, f.k0guess_x, f.k0guess_y, volumes=True)
        for w, key in ((0, "ov0"), (1, "ov1")):
            err = rel_l2(got[dd, :, w], ref[key])
            log("overlaps dir %d %s rel-L2 %.3e" % (dd, key, err))
            assert err < TOL

    al = str(tmp_path / "al.dv")
    run("al", "--savealignedraw", al)
    assert os.path.getsize(al) == 1024
    h = mrc.read_header(open(al, "rb").read(1024))
    assert (h["nx"], h["mode"]) == (64, 2)
    assert h["labels"][h["nlab"] - 1].startswith("drift-corrected raw images")


def test_cli_tiff_saveoverlaps_matches_oracle(tmp_path):
    """TIFF mode: --saveoverlaps writes one float TIFF (width 2*nx, 2*nz pages:
    overlap0's planes, then overlap1's), rewritten per direction, so the last
    direction's pair is left (cudaSirecon.cpp:330-339); no result file."""
    p, otf, raw = load_testdata_tiff()
    o = _preprocessed_oracle(p, otf, raw)
    dd = tmp_path / "data"
    dd.mkdir()
    shutil.copy(os.path.join(GOLDEN, "raw.tif"), dd / "raw.tif")
    ov = str(tmp_path / "ov.tif")
    r = run_exe([str(dd), "raw", os.path.join(GOLDEN, "otf.tif"), "-c", os.path.join(GOLDEN, "config-tiff"),
                 "--saveoverlaps", ov])
    assert "Quit processing" in r.stdout
    assert not (dd / "GPUsirecon" / "raw_proc.tif").exists()
    vol = tiff_frames(ov)
    assert vol.shape == (18, 64, 128) and vol.dtype == np.float32
    got = np.ascontiguousarray(vol).view(np.complex64)
    f = o.fit(2)
    ref = o.makeoverlaps(2, 0, 2, f.k0guess_x, f.k0guess_y, volumes=True)
    for sl, key in ((slice(0, 9), "ov0"), (slice(9, 18), "ov1")):
        err = rel_l2(got[sl], ref[key])
        log("tiff overlaps %s rel-L2 %.3e" % (key, err))
        assert err < TOL


def test_cli_full_otf_mrc_matches_oracle(tmp_path):
    """An OTF file that is not rotationally averaged (no otfRA): nzotf sections of
    [ky][kx] complex per order, dkrotf = xlen and dkzotf = zlen (cudaSirecon.cpp:1505-1515,
    1596-1603), read by trilinear lookups."""
    p, otf, raw = load_testdata()
    q, full = full_otf_from_ra(p, otf)
    q.dkrotf = float(np.float32(q.dkrotf))
    q.dkzotf = float(np.float32(q.dkzotf))
    norders, nzotf, nyotf, nxotf = full.shape
    h = mrc.make_header(nxotf, nyotf, norders * nzotf, 4, xlen=q.dkrotf, ylen=q.dkrotf, zlen=q.dkzotf)
    otf_path = str(tmp_path / "full_otf.dv")
    mrc.write(otf_path, full.reshape(norders * nzotf, nyotf, nxotf), h)
    cfg = tmp_path / "config-full"
    cfg.write_text("".join(l for l in open(os.path.join(GOLDEN, "config")) if not l.startswith("otfRA")))
    out = str(tmp_path / "full_proc.dv")
    r = run_exe([os.path.join(GOLDEN, "raw.dv"), out, otf_path, "-c", str(cfg)])
    assert "nyotf=%d" % nyotf in r.stdout
    _, vol = mrc.read(out)
    ref = oracle_series(q, full, [raw])[0]
    err = rel_l2(vol, ref)
    log("full OTF (mrc) rel-L2 %.3e" % err)
    assert err < TOL
