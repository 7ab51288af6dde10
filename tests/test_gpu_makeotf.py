"""GPU parity of the makeotf path (simr_makeotf and the makeotf executable)
against the CPU oracle restatement of src/otf/radialft.cpp.

Tolerances: OTF rel-L2 <= 1e-5 (complex), bead centre within 1e-4 px,
side-band phases within 1e-5 rad.
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from cases import GOLDEN, load_testdata_tiff, rel_l2, tiff_frames
from cudasirecon_amd import engine, mrc, synth
from cudasirecon_amd._abi import otf_default_params, set_otf_params
from oracle import binding as ob

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAKEOTF = os.path.join(ROOT, "cudasirecon_amd", "bin", "makeotf")
CUDASIRECON = os.path.join(ROOT, "cudasirecon_amd", "bin", "cudasirecon")


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def params(**kw):
    p = otf_default_params()
    set_otf_params(p, dr=0.08, dz=0.125, linespacing=0.2035, k0angle=-0.8, na=1.42, nimm=1.515, wavelength=515)
    return set_otf_params(p, **kw)


CASES = {
    # 5 phases, prime z count (direct-DFT z pass), border background, fixorigin + cleanup
    "p5_nz17": (dict(nphases=5, interpkr=[3, 20], leavekz=[7, 11, 3]), (64, 64, 17), None),
    # 3 phases, five bands, conjugate, constant background, ifixkr, per-order rescale
    "p3_5bands": (dict(nphases=3, five_bands=1, conjugate=1, background=90.0, ifixr=2, dorescale=1), (48, 48, 16), None),
    # widefield 2D, cosine apodization
    "p1_2d": (dict(nphases=1, napodize=-1, background=100.0), (64, 64, 1), None),
    # 5 phases, given per-section backgrounds, two-lens cleanup, pi shift, bead compensation
    "p5_twolens": (dict(nphases=5, twolens=1, leavekz=[5, 9, 2], forced_pi_shift=1, bead_diameter=0.17),
                   (96, 96, 20), "bg"),
}


@pytest.mark.parametrize("name", list(CASES))
def test_makeotf_matches_oracle(name):
    kw, (nx, ny, nz), bgmode = CASES[name]
    log("makeotf case", name)
    p = params(**kw)
    psf = synth.make_psf_stack(nx, ny, nz, p.nphases, seed=len(name))
    bg = None
    if bgmode == "bg":
        bg = np.random.default_rng(7).uniform(95, 105, size=psf.shape[0]).astype(np.float32)
    got, ginfo, gb0 = engine.makeotf(p, psf, section_bg=bg, want_band0=True)
    ref, rinfo, rb0 = ob.makeotf(p, psf, section_bg=bg, want_band0=True)
    assert got.shape == ref.shape
    err = rel_l2(got, ref)
    log("  otf rel-L2 %.3e  centre (%.4f %.4f %.4f)" % (err, ginfo.xc, ginfo.yc, ginfo.zc))
    assert err < 1e-5
    for a in ("xc", "yc", "zc"):
        assert abs(getattr(ginfo, a) - getattr(rinfo, a)) < 1e-4
    for o in range(1, ginfo.norders):
        assert abs(ginfo.phi[o] - rinfo.phi[o]) < 1e-5
    assert abs(ginfo.scalefactor / rinfo.scalefactor - 1) < 1e-5
    assert rel_l2(gb0, rb0) < 1e-6


def _write_tiff(path, vol):
    from PIL import Image

    ims = [Image.fromarray(np.ascontiguousarray(f, dtype=np.float32), mode="F") for f in vol]
    ims[0].save(path, save_all=True, append_images=ims[1:])


def test_makeotf_cli_tiff_and_reconstruction(tmp_path):
    """makeotf on a TIFF PSF, then cudasirecon with that OTF on the test_data TIFF."""
    psf = synth.make_psf_stack(64, 64, 32, 5, seed=11)
    src = str(tmp_path / "psf.tif")
    _write_tiff(src, psf)
    out = str(tmp_path / "otf.tif")
    args = [MAKEOTF, src, out, "-angle", "-0.8043", "-ls", "0.2035", "-na", "1.42", "-nimm", "1.515",
            "-fixorigin", "3", "20", "-xyres", "0.08", "-zres", "0.125", "-wavelength", "515"]
    log("run makeotf (tiff)")
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    pages = tiff_frames(out)
    assert pages.shape == (3, 33, 64)  # bands x kr x (2*nz floats)
    got = np.ascontiguousarray(pages).view(np.complex64)
    p = params(nphases=5, k0angle=-0.8043, interpkr=[3, 20])
    ref, _, _ = ob.makeotf(p, psf)
    assert rel_l2(got, ref) < 1e-5
    amp = tiff_frames(str(tmp_path / "otf_OnlyForViewing.tif"))
    np.testing.assert_allclose(amp, np.sqrt(np.abs(ref)), rtol=1e-4, atol=1e-6)

    # reconstruct the reference's TIFF test data with the new OTF
    d = tmp_path / "data"
    d.mkdir()
    shutil.copy(os.path.join(GOLDEN, "raw.tif"), d / "raw.tif")
    cfg = tmp_path / "config-tiff"
    cfg.write_text(open(os.path.join(GOLDEN, "config-tiff")).read())
    log("run cudasirecon with the new OTF")
    r = subprocess.run([CUDASIRECON, str(d), "raw", out, "-c", str(cfg)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "!!Error" not in r.stderr, r.stderr
    vol = tiff_frames(str(d / "GPUsirecon" / "raw_proc.tif"))
    q, _, raw = load_testdata_tiff()
    q = q.copy()
    q.nzotf, q.nxotf = 32, 33
    q.dkzotf = float(np.float32(1) / (np.float32(0.125) * np.float32(32)))
    q.dkrotf = float(np.float32(1) / (np.float32(0.08) * np.float32(32) * np.float32(2)))
    o = ob.Oracle(q, np.ascontiguousarray(ref), 32, True)
    o.load(raw)
    o.rescale(0)
    o.set_time(0)
    o.process_volume()
    err = rel_l2(vol, o.result())
    log("reconstruction with makeotf OTF rel-L2 %.3e" % err)
    assert err < 1e-5


def test_makeotf_cli_mrc(tmp_path):
    psf = synth.make_psf_stack(48, 48, 20, 5, seed=12)
    h = mrc.make_header(48, 48, psf.shape[0], 2, xlen=0.08, ylen=0.08, zlen=0.125, wave=515,
                        labels=["bead psf"])
    src = str(tmp_path / "psf.dv")
    mrc.write(src, psf, h)
    out = str(tmp_path / "otf.dv")
    log("run makeotf (mrc)")
    r = subprocess.run([MAKEOTF, src, out, "-angle", "-0.8", "-ls", "0.2035", "-na", "1.42", "-leavekz", "7", "11", "3"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ho, data = mrc.read(out)
    assert (ho["nx"], ho["ny"], ho["nz"], ho["mode"]) == (20, 25, 3, 4)
    assert ho["xlen"] == pytest.approx(1 / (20 * 0.125)) and ho["ylen"] == pytest.approx(1 / (48 * 0.08))
    # MRC input: header pixel sizes and wavelength, background 0 (no -bgInExtHdr)
    p = params(nphases=5, leavekz=[7, 11, 3])
    ref, _, _ = ob.makeotf(p, psf, section_bg=np.zeros(psf.shape[0], np.float32))
    assert rel_l2(data, ref) < 1e-5


I2M_CASES = {
    "i2m_64x16": ((64, 64, 16), dict(), 21),
    "i2m_fixorigin": ((64, 64, 12), dict(interpkr=[3, 14]), 22),
    "i2m_prime_nz": ((48, 48, 11), dict(), 23),
    "i2m_2d": ((64, 64, 1), dict(), 24),
}


@pytest.mark.parametrize("name", list(I2M_CASES))
def test_makeotf_i2m_matches_oracle(name):
    (nx, ny, nz), kw, seed = I2M_CASES[name]
    log("makeotf -I2M case", name)
    p = params(**kw)
    vol = synth.make_psf_stack(nx, ny, nz, 1, seed=seed, center=(nx / 2 - 1.3, ny / 2 + 0.6, nz / 2 - 0.3))
    got, ginfo = engine.makeotf_i2m(p, vol)
    ref, rinfo = ob.makeotf_i2m(p, vol)
    err = rel_l2(got, ref)
    log("  i2m otf rel-L2 %.3e  centre (%.4f %.4f %.4f)" % (err, ginfo.xc, ginfo.yc, ginfo.zc))
    assert err < 1e-5
    for a in ("xc", "yc", "zc"):
        assert abs(getattr(ginfo, a) - getattr(rinfo, a)) < 1e-4


def test_makeotf_cli_mrc_i2m(tmp_path):
    """-I2M: each z holds the 5 phases and then one I2M section."""
    nx = ny = 48
    nz = 14
    psf = synth.make_psf_stack(nx, ny, nz, 5, seed=31)
    i2m = synth.make_psf_stack(nx, ny, nz, 1, seed=32, center=(22.6, 25.2, 6.4), background=80.0)
    sec = np.concatenate([psf.reshape(nz, 5, ny, nx), i2m.reshape(nz, 1, ny, nx)], axis=1).reshape(nz * 6, ny, nx)
    h = mrc.make_header(nx, ny, sec.shape[0], 2, xlen=0.08, ylen=0.08, zlen=0.125, wave=515, labels=["bead psf + i2m"])
    src = str(tmp_path / "psf.dv")
    mrc.write(src, sec, h)
    out = str(tmp_path / "otf.dv")
    out_i2m = str(tmp_path / "i2m_otf.dv")
    log("run makeotf -I2M (mrc)")
    r = subprocess.run([MAKEOTF, src, out, "-angle", "-0.8", "-ls", "0.2035", "-na", "1.42", "-fixorigin", "3", "12",
                        "-I2M", out_i2m], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "I2M psf's center of mass" in r.stdout
    ho, data = mrc.read(out)
    assert (ho["nx"], ho["ny"], ho["nz"], ho["mode"]) == (nz, 25, 3, 4)
    p = params(nphases=5, interpkr=[3, 12])
    ref, _, _ = ob.makeotf(p, psf, section_bg=np.zeros(psf.shape[0], np.float32))
    assert rel_l2(data, ref) < 1e-5
    hi, di = mrc.read(out_i2m)
    assert (hi["nx"], hi["ny"], hi["nz"], hi["mode"]) == (nz, 25, 1, 4)
    assert hi["xlen"] == pytest.approx(1 / (nz * 0.125)) and hi["ylen"] == pytest.approx(1 / (nz * 0.125))
    assert hi["zlen"] == pytest.approx(1 / (nx * 0.08))
    ref_i2m, _ = ob.makeotf_i2m(p, i2m)
    err = rel_l2(np.asarray(di).reshape(25, nz), ref_i2m)
    log("  cli i2m otf rel-L2 %.3e" % err)
    assert err < 1e-5


def test_makeotf_cli_i2m_refuses_tiff(tmp_path):
    psf = synth.make_psf_stack(48, 48, 6, 6, seed=2)
    src = str(tmp_path / "psf.tif")
    _write_tiff(src, psf)
    r = subprocess.run([MAKEOTF, src, str(tmp_path / "o.tif"), "-I2M", str(tmp_path / "i.tif")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "-I2M needs MRC input" in r.stderr
