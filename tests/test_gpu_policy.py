"""GPU parity of the k0 policy branches and the flat-field correction: the HIP
path against the FP32 oracle (same precision plan: FP64 fit reductions) on
the same inputs.  Bars: output rel-L2 <= 1e-5, k0 within 1e-3 px, modamp
within 1e-3.

  * --usecorr flat-field maps (bgAndSlope / load_and_flatfield,
    cudaSirecon.cpp:139-192, 1715-1731), through the C ABI and through the
    executable with an MRC calibration file;
  * forcemodamp (cudaSirecon.cpp:459-473): modamps below the forced value are
    raised to it;
  * a time series searched at every time point (k0searchAll) whose second
    frame's pattern lies more than 0.1 1/um from the guess: the fitted k0 is
    dropped for time 0's and the modamps are refitted there (380-448);
  * 2D with five phases: order 2's modamp is chained, amp[2] *= amp[1]
    (450-457).
"""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from cases import GOLDEN, load_testdata, rel_l2, synthetic
from cudasirecon_amd import engine, mrc, synth
from oracle import binding as ob

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cudasirecon_amd", "bin", "cudasirecon")


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fit_errors(p, fa, fb):
    dk = 1.0 / (p.nx * p.dxy)
    norders = engine.norders_of(p)
    k0e = max(math.hypot(a.k0x - b.k0x, a.k0y - b.k0y) / dk for a, b in zip(fa, fb))
    ame = max(abs(complex(a.amp_re[o], a.amp_im[o]) - complex(b.amp_re[o], b.amp_im[o]))
              for a, b in zip(fa, fb) for o in range(norders))
    return k0e, ame


def series(p, otf, frames, maps=None):
    """(GPU volumes, GPU fits, oracle volumes, oracle fits) of one sequential context each."""
    gv, gf, ov, of = [], [], [], []
    o = ob.Oracle(p, otf, 32, True)
    if maps is not None:
        o.set_flatfield(*maps)
    with engine.Reconstructor(p, otf) as g:
        if maps is not None:
            g.set_flatfield(*maps)
        for t, raw in enumerate(frames):
            g.set_raw(raw)
            g.load_and_rescale(t)
            g.set_time(t)
            g.process_volume()
            gv.append(g.result())
            gf.append([g.fit(d) for d in range(p.ndirs)])
            o.load(raw.astype(np.float32))
            o.rescale(t)
            o.set_time(t)
            o.process_volume()
            ov.append(o.result())
            of.append([o.fit(d) for d in range(p.ndirs)])
    return gv, gf, ov, of


def check(name, p, gv, gf, ov, of):
    for t in range(len(gv)):
        k0e, ame = fit_errors(p, gf[t], of[t])
        e = rel_l2(gv[t], ov[t])
        log("  %s t=%d: rel-L2 %.2e, k0 %.2e px, modamp %.2e" % (name, t, e, k0e, ame))
        assert k0e < 1e-3 and ame < 1e-3, (name, t, k0e, ame)
        assert e < 1e-5, (name, t, e)


def test_flatfield_maps_match_oracle():
    log("case: 64x64x8 recorded through a camera with per-pixel offset and gain, --usecorr maps")
    p, otf, raw, _ = synthetic(64, 64, 8, seed=21)
    rng = np.random.default_rng(4)
    b = rng.uniform(80.0, 120.0, (p.ny, p.nx)).astype(np.float32)
    s = rng.uniform(0.8, 1.25, (p.ny, p.nx)).astype(np.float32)
    recorded = (b + (raw.astype(np.float64) - p.constbkgd) / s).astype(np.float32)
    q = p.copy()
    q.constbkgd = 0.0
    gv, gf, ov, of = series(q, otf, [recorded], maps=(b, s))
    check("usecorr", q, gv, gf, ov, of)


def test_cli_usecorr_matches_oracle(tmp_path):
    log("case: executable on test_data raw.dv with an MRC calibration file (--usecorr)")
    p, otf, raw = load_testdata()
    rng = np.random.default_rng(9)
    b = rng.uniform(-5.0, 5.0, (p.ny, p.nx)).astype(np.float32)
    s = rng.uniform(0.9, 1.1, (p.ny, p.nx)).astype(np.float32)
    corr = str(tmp_path / "corr.mrc")
    mrc.write(corr, np.stack([b, s]), mrc.make_header(p.nx, p.ny, 2, 2))
    out = str(tmp_path / "proc.dv")
    r = subprocess.run([EXE, os.path.join(GOLDEN, "raw.dv"), out, os.path.join(GOLDEN, "otf.dv"), "-c",
                        os.path.join(GOLDEN, "config"), "--usecorr", corr], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and "!!Error occurred" not in r.stderr, r.stderr
    assert "loading CCD calibration file" in r.stdout
    _, vol = mrc.read(out)
    o = ob.Oracle(p, otf, 32, True)
    o.set_flatfield(b, s)
    o.load(raw)
    o.rescale(0)
    o.set_time(0)
    o.process_volume()
    e = rel_l2(vol, o.result())
    log("  cli usecorr rel-L2 %.2e" % e)
    assert e < 1e-5


def test_forcemodamp_matches_oracle():
    log("case: 64x64x8 with forcemodamp 0.9 0.9 (raises both orders' modamp magnitudes)")
    p, otf, raw, _ = synthetic(64, 64, 8, seed=22)
    q = p.copy()
    q.nforceamp, q.forceamp[0], q.forceamp[1] = 2, 0.9, 0.9
    gv, gf, ov, of = series(q, otf, [raw])
    check("forcemodamp", q, gv, gf, ov, of)
    for f in gf[0]:
        for o in (1, 2):
            assert abs(complex(f.amp_re[o], f.amp_im[o])) >= 0.9 - 1e-5


def test_series_search_falls_back_to_time0_k0():
    log("case: 3-frame 64x64x8 series searched at every frame; frame 1's pattern rotated by 4 degrees "
        "(more than 0.1 1/um from the guess), frame 2 jittered by 0.05 px")
    angles = [1.648 + i * math.pi / 3 for i in range(3)]
    p, otf, raw0, _ = synth.make_dataset(64, 64, 8, seed=23, angles=angles)
    _, _, raw1, _ = synth.make_dataset(64, 64, 8, seed=24, angles=[a + math.radians(4.0) for a in angles])
    _, _, raw2, _ = synth.make_dataset(64, 64, 8, seed=25, angles=angles, k0_jitter_px=0.05)
    q = p.copy()
    q.ntimes, q.bUseTime0k0 = 3, 0
    gv, gf, ov, of = series(q, otf, [raw0, raw1, raw2])
    check("k0 search per frame", q, gv, gf, ov, of)
    for fits in (gf, of):
        # frame 1 took time 0's k0; frame 2 kept its own fit
        assert all((fits[1][d].k0x, fits[1][d].k0y) == (fits[0][d].k0x, fits[0][d].k0y) for d in range(3))
        assert any((fits[2][d].k0x, fits[2][d].k0y) != (fits[0][d].k0x, fits[0][d].k0y) for d in range(3))


def test_2d_five_phases_chains_order2_modamp():
    log("case: 2D 128x128, 3 dirs x 5 phases (order 2 modamp chained on order 1)")
    p, otf, raw, _ = synthetic(128, 128, 1, nphases=5, seed=26)
    gv, gf, ov, of = series(p, otf, [raw])
    check("2D nphases 5", p, gv, gf, ov, of)


def _run(p, otf, raw, env):
    os.environ.update(env)
    try:
        with engine.Reconstructor(p, otf) as g:
            g.set_raw(raw)
            g.load_and_rescale(0)
            g.set_time(0)
            g.process_volume()
            return g.result(), [g.fit(d) for d in range(p.ndirs)]
    finally:
        for k in env:
            os.environ.pop(k, None)


@pytest.mark.parametrize("case", ["3d_fp64", "3d_fp32", "2d_3ph", "2d_5ph", "testdata", "fallback"])
def test_device_fit_matches_host_fit(case):
    """The line searches driven on the device (the default) against the same
    walks replayed on the host with a read-back per round (SIMR_FIT_ON_HOST=1):
    the same candidates are evaluated in the same order (so the evaluation
    counts agree), and k0, modamps and the output agree bit for bit (the
    angles and candidate vectors come from the same reproducible cos, sin and
    atan2 on both sides).  "fallback": one device round only
    (SIMR_FIT_ROUNDS=1), so the device fit stops short and the host walk
    takes over."""
    log("device fit vs host fit:", case)
    if case in ("testdata", "fallback"):
        p, otf, raw = load_testdata()
    else:
        nz = 1 if case.startswith("2d") else 12
        nph = 5 if case in ("3d_fp64", "3d_fp32", "2d_5ph") else 3
        p, otf, raw, _ = synthetic(96, 96, nz, nphases=nph)
        if case == "3d_fp32":
            p.fit_fp64 = 0
    vd, fd = _run(p, otf, raw, {"SIMR_FIT_ROUNDS": "1"} if case == "fallback" else {})
    vh, fh = _run(p, otf, raw, {"SIMR_FIT_ON_HOST": "1"})
    k0e, ame = fit_errors(p, fd, fh)
    e = rel_l2(vd, vh)
    log("  rel-L2 %.2e, k0 %.2e px, modamp %.2e" % (e, k0e, ame))
    for a, b in zip(fd, fh):
        assert (a.n_angle_evals, a.n_mag_evals, a.n_reductions) == (b.n_angle_evals, b.n_mag_evals, b.n_reductions)
    assert k0e == 0.0 and ame == 0.0, (k0e, ame)
    assert e == 0.0, e


@pytest.mark.parametrize("case", ["3d_512", "3d_256", "2d_256", "testdata"])
def test_two_pass_overlaps_match_three_pass(case):
    """The FP64 fit's overlaps built in two passes (formation with the y
    transform by columns, then the x transform with the z sums by rows)
    against the three-pass form (SIMR_OVERLAPS_3PASS=1): the 2D inverse
    transform is taken in the other order, so P differs by rounding only, and
    the fitted k0, the modamps and the output agree to that level."""
    log("two-pass vs three-pass overlaps:", case)
    if case == "testdata":
        p, otf, raw = load_testdata()
    elif case == "2d_256":
        p, otf, raw, _ = synthetic(256, 256, 1, nphases=3)
    else:
        n = 512 if case == "3d_512" else 256
        p, otf, raw, _ = synthetic(n, n, 16, nphases=5)
    v2, f2 = _run(p, otf, raw, {})
    v3, f3 = _run(p, otf, raw, {"SIMR_OVERLAPS_3PASS": "1"})
    k0e, ame = fit_errors(p, f2, f3)
    e = rel_l2(v2, v3)
    log("  rel-L2 %.2e, k0 %.2e px, modamp %.2e" % (e, k0e, ame))
    assert k0e < 1e-4 and ame < 1e-5, (k0e, ame)
    assert e < 2e-6, e
