"""Option handling of the drop-in cudasirecon command line and config files (CPU).

Each case states what the reference's option table and setParams do
(src/cudaSirecon/cudaSirecon.cpp:16-72, 913-1071) under Boost.Program_options'
store/notify semantics, and checks SR_parseArgs / SR_parseConfig
(include/cudasirecon.h), which run the same parser as SR_new_from_argv and the
cudasirecon executable without touching data files or a GPU.
"""
import ctypes as C
import os
import subprocess

import pytest

from cases import GOLDEN
from cudasirecon_amd import mrc, sr
from cudasirecon_amd.engine import SimrError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cudasirecon_amd", "bin", "cudasirecon")


def parse(*args):
    return sr.parse_args(["cudasirecon"] + list(args))


def write_cfg(tmp_path, text, name="cfg"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_option_defaults_override_setdefaultparams():
    rc, r, im = parse("in", "out", "otf")
    assert rc == 0
    # option defaults (cudaSirecon.cpp:918-1005) replace SetDefaultParams' values
    assert (r.ndirs, r.nphases, r.norders_output) == (3, 5, 0)
    assert r.k0startangle == pytest.approx(1.648) and r.linespacing == pytest.approx(0.172)
    assert r.na == pytest.approx(1.2) and r.nimm == pytest.approx(1.33)
    assert r.wiener == pytest.approx(0.01) and r.otfcutoff == pytest.approx(0.006)
    assert r.zoomfact == 2 and r.z_zoom == 1 and r.constbkgd == 0
    assert r.BesselNA == pytest.approx(0.144) and r.BesselLambdaEx == pytest.approx(0.488)
    assert r.dzPSF == pytest.approx(0.15)
    assert im.dxy == pytest.approx(0.1) and im.dz == pytest.approx(0.2) and im.wave[0] == 530
    # fields without an option keep SetDefaultParams' values
    assert (r.napodize, r.suppression_radius, r.recalcarrays, r.bFitallphases) == (10, 10, 1, 1)
    assert (r.bSearchforvector, r.bUseTime0k0, r.do_rescale, r.bSuppress_singularities) == (1, 1, 1, 1)
    assert r.bFilteroverlaps == 1 and r.nforceamp == 1 and r.forceamp[0] == 0
    # gammaApo always has a value, so the output apodization is always mode 2
    assert r.apodizeoutput == 2 and r.apoGamma == 1
    # positionals
    assert (r.ifiles, r.ofiles, r.otffiles) == (b"in", b"out", b"otf")
    assert r.bRadAvgOTF == 0 and r.bTIFF == 0


def test_bare_implicit_flags_store_zero():
    _, r, _ = parse("in", "out", "otf", "--k0searchAll", "--norescale", "--nosuppress", "--nofilterovlps")
    assert (r.bUseTime0k0, r.do_rescale, r.bSuppress_singularities, r.bFilteroverlaps) == (0, 0, 0, 0)
    _, r, _ = parse("in", "out", "otf", "--k0searchAll=1", "--norescale=0")
    assert (r.bUseTime0k0, r.do_rescale) == (1, 0)


def test_config_file_booleans_are_inverted(tmp_path):
    # in a config file the value is stored as given: norescale=1 keeps rescaling,
    # k0searchAll=1 keeps time-0 k0 reuse, nosuppress=1 keeps suppression
    cfg = write_cfg(tmp_path, "norescale=1\nk0searchAll=1\nnosuppress=1\nnofilterovlps=1\n")
    _, r, _ = parse("in", "out", "otf", "-c", cfg)
    assert (r.do_rescale, r.bUseTime0k0, r.bSuppress_singularities, r.bFilteroverlaps) == (1, 1, 1, 1)
    cfg = write_cfg(tmp_path, "norescale=0\nk0searchAll=0\nnofilterovlps=false\n", "cfg2")
    _, r, _ = parse("in", "out", "otf", "--config", cfg)
    assert (r.do_rescale, r.bUseTime0k0, r.bFilteroverlaps) == (0, 0, 0)


def test_switches_on_cli_and_in_config(tmp_path):
    _, r, _ = parse("in", "out", "otf", "--otfRA", "--fastSI", "--dampenOrder0", "--equalizez", "--nokz0")
    assert (r.bRadAvgOTF, r.bFastSIM, r.bDampenOrder0, r.equalizez, r.bNoKz0, r.equalizet) == (1, 1, 1, 1, 1, 0)
    cfg = write_cfg(tmp_path, "otfRA=1\nfastSI=yes\ndampenOrder0=0\notfPerAngle=true\n")
    _, r, _ = parse("in", "out", "otf", "-c", cfg)
    assert (r.bRadAvgOTF, r.bFastSIM, r.bDampenOrder0, r.bOneOTFperAngle) == (1, 1, 0, 1)
    with pytest.raises(SimrError, match="does not take any arguments"):
        parse("in", "out", "otf", "--otfRA=1")


def test_command_line_wins_over_config(tmp_path):
    cfg = write_cfg(tmp_path, "wiener=0.5\nndirs=2\n# comment\n  ls = 0.25  \n")
    _, r, _ = parse("in", "out", "otf", "--wiener", "0.002", "-c", cfg)
    assert r.wiener == pytest.approx(0.002)
    assert r.ndirs == 2 and r.linespacing == pytest.approx(0.25)


def test_golden_configs():
    _, r, im = parse("raw.dv", "proc.dv", "otf.dv", "-c", os.path.join(GOLDEN, "config"))
    assert r.nimm == pytest.approx(1.515) and r.wiener == pytest.approx(0.001) and r.na == pytest.approx(1.42)
    assert r.linespacing == pytest.approx(0.2035) and r.bRadAvgOTF == 1 and r.bDampenOrder0 == 1
    assert r.nk0angles == 3
    assert [r.k0angles[i] for i in range(3)] == pytest.approx([-0.8043, -1.8555, 0.2388])
    r, im = sr.parse_config(os.path.join(GOLDEN, "config-tiff"))
    assert r.bFastSIM == 1 and im.dxy == pytest.approx(0.08) and im.dz == pytest.approx(0.125)
    assert r.dzPSF == pytest.approx(0.125)
    assert [r.k0angles[i] for i in range(3)] == pytest.approx([0.8043, 1.8555, -0.2388])


def test_list_options(tmp_path):
    _, r, _ = parse("in", "out", "otf", "--forcemodamp", "0.5", "0.25", "--k0angles", "1.0,,-2.5,3")
    assert r.nforceamp == 2 and list(r.forceamp[:2]) == pytest.approx([0.5, 0.25])
    assert r.nk0angles == 3 and list(r.k0angles[:3]) == pytest.approx([1.0, -2.5, 3.0])
    _, r, _ = parse("--forcemodamp", "0.7", "--angle0", "-0.8", "in", "out", "otf")
    assert r.nforceamp == 1 and r.forceamp[0] == pytest.approx(0.7) and r.k0startangle == pytest.approx(-0.8)


def test_prefix_guessing_and_errors(tmp_path):
    _, r, _ = parse("in", "out", "otf", "--wien", "0.003", "--nph", "3")
    assert r.wiener == pytest.approx(0.003) and r.nphases == 3
    with pytest.raises(SimrError, match="ambiguous"):
        parse("in", "out", "otf", "--otf", "x")
    with pytest.raises(SimrError, match="unrecognised option '--bogus'"):
        parse("in", "out", "otf", "--bogus", "1")
    with pytest.raises(SimrError, match="required but missing"):
        parse("in", "out")
    with pytest.raises(SimrError, match="is invalid"):
        parse("in", "out", "otf", "--ndirs", "two")
    with pytest.raises(SimrError, match="is invalid"):
        parse("in", "out", "otf", "--ndirs", "3.5")
    with pytest.raises(SimrError, match="cannot be specified more than once"):
        parse("in", "out", "otf", "--ndirs", "3", "--ndirs", "2")
    with pytest.raises(SimrError, match="too many positional"):
        parse("in", "out", "otf", "extra")
    with pytest.raises(SimrError, match="missing"):
        parse("in", "out", "otf", "--wiener")
    bad = write_cfg(tmp_path, "nonsense=1\n", "bad1")
    with pytest.raises(SimrError, match="unrecognised option 'nonsense'"):
        parse("in", "out", "otf", "-c", bad)
    bad = write_cfg(tmp_path, "wiener 0.1\n", "bad2")
    with pytest.raises(SimrError, match="invalid line"):
        parse("in", "out", "otf", "-c", bad)


def test_missing_config_file_proceeds():
    rc, r, _ = parse("in", "out", "otf", "-c", "/nonexistent/cfg")
    assert rc == 0 and r.wiener == pytest.approx(0.01)


def test_help_and_version():
    assert parse()[0] == 1
    assert parse("--help")[0] == 1
    assert parse("-h", "in")[0] == 1
    assert parse("--version")[0] == 2


def test_struct_layout_matches_c(tmp_path):
    """ctypes mirrors against offsetof() of the C header, compiled with gcc."""
    checks = []
    for cls, cname in ((sr.SRReconParams, "SR_ReconParams"), (sr.SRImageParams, "SR_ImageParams")):
        checks.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for name, _ in cls._fields_:
            checks.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, name, cname, name))
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "cudasirecon.h"\nint main(void){%s return 0;}\n'
                   % "\n".join(checks))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    for cls, cname in ((sr.SRReconParams, "SR_ReconParams"), (sr.SRImageParams, "SR_ImageParams")):
        assert int(got[cname]) == C.sizeof(cls)
        for name, _ in cls._fields_:
            assert int(got["%s.%s" % (cname, name)]) == getattr(cls, name).offset, name


def test_executable_help_and_error_exit_status(tmp_path):
    r = subprocess.run([EXE, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "Usage:" in r.stdout and "--otfRA" in r.stdout
    r = subprocess.run([EXE, "--version"], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("Version")
    # errors are printed and, as in the reference driver, the exit status stays 0
    r = subprocess.run([EXE, "in.dv", "out.dv"], capture_output=True, text=True)
    assert r.returncode == 0 and "!!Error occurred: the option '--otf-file' is required but missing" in r.stderr
    r = subprocess.run([EXE, str(tmp_path / "missing.dv"), str(tmp_path / "o.dv"), "otf.dv", "--otfRA"],
                       capture_output=True, text=True)
    assert "!!Error occurred: No matching input TIFF or MRC files not found" in r.stderr


def test_deskew_to_an_unsupported_width_is_refused(tmp_path):
    # raw.dv: nx_raw 64, 9 sections per phase, dz 0.2 / dxy 0.1 -> findOptimalDimension(64 + 9 cos(30) 2) = 75,
    # a width the engine has no kernels for: a clear error, no partial result
    r = subprocess.run([EXE, os.path.join(GOLDEN, "raw.dv"), str(tmp_path / "o.dv"), os.path.join(GOLDEN, "otf.dv"),
                        "--deskew", "30", "--otfRA"], capture_output=True, text=True)
    assert "!!Error occurred: image width 75 is not a supported size" in r.stderr, r.stderr
    assert "nx=75" in r.stdout


def test_unsupported_branches_are_refused(tmp_path):
    for extra in (["--bessel", "--noRecon"], ["--2lenses"]):
        r = subprocess.run([EXE, os.path.join(GOLDEN, "raw.dv"), str(tmp_path / "o.dv"),
                            os.path.join(GOLDEN, "otf.dv")] + extra + ["--otfRA"],
                           capture_output=True, text=True)
        assert "!!Error occurred:" in r.stderr and "not supported" in r.stderr, (extra, r.stderr)


def test_ra_otf_file_without_otfra_is_refused(tmp_path):
    # without --otfRA the file is read as a full 3D OTF (cudaSirecon.cpp:1505-1515): the RA
    # file's zlen of 0 gives no kz step, which is an error here, not a reconstruction
    r = subprocess.run([EXE, os.path.join(GOLDEN, "raw.dv"), str(tmp_path / "o.dv"), os.path.join(GOLDEN, "otf.dv")],
                       capture_output=True, text=True)
    assert "!!Error occurred: the OTF file does not hold a full 3D OTF" in r.stderr, r.stderr
    assert "pass --otfRA" in r.stderr


def test_tiff_folder_discovery_creates_output_folder(tmp_path):
    d = tmp_path / "data"
    d.mkdir()
    for n in ("cell_t0.tif", "cell_t1.TIF", "other.tif", "cell_notes.txt"):
        (d / n).write_bytes(open(os.path.join(GOLDEN, "raw.tif"), "rb").read() if n.endswith(("tif", "TIF")) else b"x")
    # without a GPU the run stops at the device, after discovery and header reads
    r = subprocess.run([EXE, str(d), "cell", os.path.join(GOLDEN, "otf.tif"), "-c",
                        os.path.join(GOLDEN, "config-tiff")], capture_output=True, text=True)
    assert (d / "GPUsirecon").is_dir()
    assert "nx_raw=64, ny=64, nz=9" in r.stdout


def test_debug_dump_files_get_the_reference_headers(tmp_path):
    """--saveprefiltered / --saveoverlaps / --savealignedraw open their MRC files at
    setup with the input header changed as setup_common does
    (src/cudaSirecon/cudaSirecon.cpp:1179-1209).  Without a GPU the run stops at
    the device, after setup."""
    sep, ov, al = (str(tmp_path / n) for n in ("sep.dv", "ov.dv", "al.dv"))
    r = subprocess.run([EXE, os.path.join(GOLDEN, "raw.dv"), str(tmp_path / "o.dv"), os.path.join(GOLDEN, "otf.dv"),
                        "-c", os.path.join(GOLDEN, "config"), "--saveprefiltered", sep, "--saveoverlaps", ov,
                        "--savealignedraw", al], capture_output=True, text=True)
    assert "nx_raw=64, ny=64, nz=9" in r.stdout, r.stdout + r.stderr
    hin = mrc.read_header(open(os.path.join(GOLDEN, "raw.dv"), "rb").read(1024))
    hs = mrc.read_header(open(sep, "rb").read(1024))
    assert (hs["nx"], hs["ny"], hs["nz"], hs["mode"], hs["next"]) == (33, 64, hin["nz"], 4, 0)
    ho = mrc.read_header(open(ov, "rb").read(1024))
    assert (ho["nx"], ho["ny"], ho["nz"], ho["mode"]) == (64, 64, 9 * 2 * 3, 4)
    assert (ho["num_waves"], ho["interleaved"]) == (2, 1)
    ha = mrc.read_header(open(al, "rb").read(1024))
    assert (ha["nx"], ha["ny"], ha["nz"], ha["mode"]) == (64, 64, hin["nz"], 2)
    for h in (hs, ho, ha):
        assert (h["xlen"], h["zlen"], h["iwav1"]) == (hin["xlen"], hin["zlen"], hin["iwav1"])
