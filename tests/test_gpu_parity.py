"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the same inputs.

Stage by stage (same bands / same k0 and modamps in, compare out) and end to end.
Tolerances: FFT rel-L2 2e-6; stages rel-L2 1e-5; end-to-end output rel-L2 1e-5
against the FP32 oracle, k0 within 1e-3 px and modamp within 1e-3.
"""
import math
import sys

import numpy as np
import pytest

from cases import full_otf_from_ra, load_testdata, rel_l2, synthetic
from cudasirecon_amd import engine
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# every two-pass plan (4, 8, 12, 16, 20, 24, 32, 40, 48, 60, 64, 80, 96, 128, 192, 256, 512, 1024) and the
# radix-stage lengths, as rows, columns (y) and z lines
FFT_SHAPES = [(64,), (48,), (9,), (512,), (1024,), (2048,), (384,), (640,), (1280,), (4096,),
              (4,), (8,), (12,), (20,), (24,), (60,), (80,), (128,), (192,), (256,),
              (16, 8), (64, 64), (96, 128), (40, 24), (1024, 16), (512, 12), (10, 96, 128), (9, 64, 64),
              (48, 32, 16), (12, 20, 8), (60, 4, 32),
              # lengths with factors of 7 (radix-stage core)
              (7,), (14,), (21,), (49,), (56,), (112,), (196,), (224,), (28, 16), (14, 64, 64), (35, 16, 8),
              (63, 8, 8), (210, 4, 4)]


@pytest.mark.parametrize("shape", FFT_SHAPES)
def test_fft_matches_numpy(shape):
    log("fft", shape)
    rng = np.random.default_rng(len(shape) * 1000 + shape[-1])
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    for inv in (False, True):
        got = engine.fft3d(x, inverse=inv)
        ref = np.fft.ifftn(x.astype(np.complex128)) * x.size if inv else np.fft.fftn(x.astype(np.complex128))
        assert rel_l2(got, ref) < 2e-6, (shape, inv, rel_l2(got, ref))


def _gpu(p, otf, fit64=True):
    q = p.copy()
    q.fit_fp64 = 1 if fit64 else 0
    return engine.Reconstructor(q, otf)


def _oracle_preprocessed(p, otf, raw, prec=32, fit64=False):
    o = ob.Oracle(p, otf, prec, fit64)
    o.load(raw)
    o.rescale(0)
    o.set_time(0)
    o.preprocess_bands()
    return o


def _cases():
    return ["testdata", "syn3d", "syn2d", "bessel3d", "opts3d", "opts2d", "z14", "z21zz", "z11", "z13zz", "z61",
            "z131zz", "full3d", "full2d", "w2048", "w1536_2d"]


def _case(name):
    if name == "testdata":
        p, otf, raw = load_testdata()
        return p, otf, raw
    if name == "syn3d":
        p, otf, raw, _ = synthetic(128, 128, 16, nphases=5)
        return p, otf, raw
    if name == "bessel3d":  # --bessel: order-1 fit, Bessel axial cutoffs 15/14/10, anisotropic apodization
        p, otf, raw, _ = synthetic(64, 64, 32, nphases=5, dz=0.1, bBessel=1, BesselNA=0.65, BesselLambdaEx=0.488)
        return p, otf, raw
    if name == "opts3d":  # non-square, zoom 1.5, z_zoom 2, cosine apodization, no suppression, noKz0, equalizez
        p, otf, raw, _ = synthetic(96, 128, 12, nphases=5, zoomfact=1.5, z_zoom=2, apodizeoutput=1,
                                   bSuppress_singularities=0, bNoKz0=1, equalizez=1, wiener=0.01, napodize=6)
        return p, otf, raw
    if name == "z14":  # axial length with a factor 7
        p, otf, raw, _ = synthetic(64, 64, 14, nphases=5)
        return p, otf, raw
    if name == "z21zz":  # nz 21, z_zoom 2: assembly depth 42
        p, otf, raw, _ = synthetic(48, 64, 21, nphases=5, z_zoom=2)
        return p, otf, raw
    if name == "z11":  # prime depth: the direct-DFT z passes
        p, otf, raw, _ = synthetic(64, 64, 11, nphases=5)
        return p, otf, raw
    if name == "z13zz":  # prime depth 13 with z_zoom 2: assembly depth 26, both z passes by direct DFT
        p, otf, raw, _ = synthetic(48, 48, 13, nphases=5, z_zoom=2)
        return p, otf, raw
    if name == "z61":  # prime depth above 32: both z passes by direct DFT
        p, otf, raw, _ = synthetic(64, 64, 61, nphases=5)
        return p, otf, raw
    if name == "z131zz":  # prime depth 131 with z_zoom 2: assembly depth 262, above the transform kernels' 256
        p, otf, raw, _ = synthetic(32, 32, 131, nphases=5, z_zoom=2)
        return p, otf, raw
    if name == "w2048":  # a 2048-wide camera: 4096-point assembly rows, 128-thread row transforms
        p, otf, raw, _ = synthetic(2048, 48, 6, nphases=5)
        return p, otf, raw
    if name == "w1536_2d":  # 1536-wide 2D
        p, otf, raw, _ = synthetic(1536, 64, 1, nphases=3)
        return p, otf, raw
    if name == "full3d":  # a full (not rotationally averaged) 3D OTF table, trilinear lookups
        p, otf, raw, _ = synthetic(64, 64, 16, nphases=5)
        p, otf = full_otf_from_ra(p, otf)
        return p, otf, raw
    if name == "full2d":  # a full 2D OTF table [ky][kx]
        p, otf, raw, _ = synthetic(96, 96, 1, nphases=3)
        p, otf = full_otf_from_ra(p, otf)
        return p, otf, raw
    if name == "opts2d":  # non-square 2D, gamma apodization 0.8, suppression radius 5
        p, otf, raw, _ = synthetic(120, 100, 1, nphases=3, apodizeoutput=2, apoGamma=0.8, suppression_radius=5,
                                   wiener=0.005)
        return p, otf, raw
    p, otf, raw, _ = synthetic(128, 128, 1, nphases=3)
    return p, otf, raw


@pytest.mark.parametrize("case", _cases())
def test_forward_bands_match_oracle(case):
    log("bands", case)
    p, otf, raw = _case(case)
    o = _oracle_preprocessed(p, otf, raw.astype(np.float32))
    with _gpu(p, otf) as g:
        g.set_raw(raw)
        g.load_and_rescale(0)
        for d in range(p.ndirs):
            for b in range(p.nphases):
                e = rel_l2(g.band(d, b), o.band(d, b))
                assert e < 1e-5, (case, d, b, e)


@pytest.mark.parametrize("case", _cases())
def test_overlap_products_match_oracle(case):
    log("overlaps", case)
    p, otf, raw = _case(case)
    o = _oracle_preprocessed(p, otf, raw.astype(np.float32))
    fo2 = 2 if p.nz > 1 else 1
    with _gpu(p, otf) as g:
        for d in range(p.ndirs):
            for b in range(p.nphases):
                g.set_band(d, b, o.band(d, b))
        f = o.fit(0)
        for (o1, o2) in [(0, fo2), (0, 1)]:
            ref = o.makeoverlaps(0, o1, o2, f.k0guess_x, f.k0guess_y)
            got = g.debug_overlaps(0, o1, o2, f.k0guess_x, f.k0guess_y)
            for key in ("P", "sx", "sy"):
                e = rel_l2(got[key], ref[key])
                assert e < 1e-5, (case, o1, o2, key, e)


@pytest.mark.parametrize("case", _cases())
def test_overlap_volumes_match_oracle(case):
    """The overlap pair in real space (the --saveoverlaps dump): the GPU's
    stored kz planes transformed along z against the oracle's makeoverlaps
    volumes after its inverse 3D FFT."""
    log("overlap volumes", case)
    p, otf, raw = _case(case)
    o = _oracle_preprocessed(p, otf, raw.astype(np.float32))
    fo2 = 2 if p.nz > 1 and not p.bBessel else 1
    with _gpu(p, otf) as g:
        for d in range(p.ndirs):
            for b in range(p.nphases):
                g.set_band(d, b, o.band(d, b))
        d = p.ndirs - 1
        f = o.fit(d)
        for (o1, o2) in sorted({(0, fo2), (0, 1)}):
            ref = o.makeoverlaps(d, o1, o2, f.k0guess_x, f.k0guess_y, volumes=True)
            v0, v1 = g.overlap_volumes(d, o1, o2, f.k0guess_x, f.k0guess_y)
            for got, key in ((v0, "ov0"), (v1, "ov1")):
                e = rel_l2(got, ref[key])
                assert e < 1e-5, (case, o1, o2, key, e)
            assert np.abs(v0).max() > 0 and np.abs(v1).max() > 0


def _fitted_oracle(p, otf, raw):
    o = _oracle_preprocessed(p, otf, raw.astype(np.float32), 32, True)
    for d in range(p.ndirs):
        o.modulation(d)
    return o


@pytest.mark.parametrize("case", _cases())
def test_filter_matches_oracle(case):
    log("filter", case)
    p, otf, raw = _case(case)
    o = _fitted_oracle(p, otf, raw)
    with _gpu(p, otf) as g:
        for d in range(p.ndirs):
            for b in range(p.nphases):
                g.set_band(d, b, o.band(d, b))
            g.set_fit(d, o.fit(d))
        for d in range(p.ndirs):
            o.filter_dir(d)
            g.filter_dir(d)
            for b in range(p.nphases):
                e = rel_l2(g.band(d, b), o.band(d, b))
                assert e < 1e-5, (case, d, b, e)


@pytest.mark.parametrize("case", _cases())
def test_assembly_matches_oracle(case):
    log("assembly", case)
    p, otf, raw = _case(case)
    o = _fitted_oracle(p, otf, raw)
    with _gpu(p, otf) as g:
        for d in range(p.ndirs):
            for b in range(p.nphases):
                g.set_band(d, b, o.band(d, b))
            g.set_fit(d, o.fit(d))
        o.filter_assemble()
        g.filter_assemble()
        e = rel_l2(g.result(), o.result())
        assert e < 1e-5, (case, e)


def _fit_errors(p, fa, fb):
    dk = 1.0 / (p.nx * p.dxy)
    norders = p.nphases // 2 + 1
    k0e = max(math.hypot(a.k0x - b.k0x, a.k0y - b.k0y) / dk for a, b in zip(fa, fb))
    ame = max(abs(complex(a.amp_re[o], a.amp_im[o]) - complex(b.amp_re[o], b.amp_im[o]))
              for a, b in zip(fa, fb) for o in range(norders))
    return k0e, ame


@pytest.mark.parametrize("case", _cases())
def test_end_to_end_matches_oracle(case):
    log("end-to-end", case)
    p, otf, raw = _case(case)
    raw32 = raw.astype(np.float32)
    # FP64 fit on the GPU vs the oracle carrying the same precision plan
    ref, fref = ob.reconstruct(p, otf, raw32, 32, True)
    with _gpu(p, otf, fit64=True) as g:
        g.set_raw(raw)
        g.load_and_rescale(0)
        g.set_time(0)
        g.process_volume()
        vol = g.result()
        fits = [g.fit(d) for d in range(p.ndirs)]
    k0e, ame = _fit_errors(p, fits, fref)
    e = rel_l2(vol, ref)
    ref64, f64 = ob.reconstruct(p, otf, raw32, 64)
    e64 = rel_l2(vol, ref64)
    log("  fp64-fit: k0 err %.2e px, modamp err %.2e, rel-L2 vs FP32 oracle %.2e, vs FP64 oracle %.2e"
        % (k0e, ame, e, e64))
    assert k0e < 1e-3 and ame < 1e-3, (k0e, ame)
    assert e < 1e-5, e
    # FP32 fit (reference's float reductions) vs the reference-exact FP32 oracle
    ref32, f32 = ob.reconstruct(p, otf, raw32, 32, False)
    with _gpu(p, otf, fit64=False) as g:
        g.set_raw(raw)
        g.load_and_rescale(0)
        g.set_time(0)
        g.process_volume()
        vol32 = g.result()
        fits32 = [g.fit(d) for d in range(p.ndirs)]
    k0e, ame = _fit_errors(p, fits32, f32)
    e = rel_l2(vol32, ref32)
    log("  fp32-fit: k0 err %.2e px, modamp err %.2e, rel-L2 vs FP32 oracle %.2e" % (k0e, ame, e))
    assert k0e < 1e-3 and ame < 1e-3, (k0e, ame)
    assert e < 1e-5, e


def test_nan_pixel_reports_error_and_context_recovers():
    """One NaN pixel in a float stack spreads through the transforms to every
    cross-correlation value.  The peak search then keeps findpeak's initial
    centre (gpuFunctionsImpl.cu:443-455) - the oracle's k0 turns NaN there -
    and the engine reports the non-finite k0 as an error instead of taking it
    into the overlap and filter kernels.  The same context then reconstructs a
    clean stack as the oracle does."""
    log("case: testdata with raw[7, 20, 30] = NaN, then the clean stack")
    p, otf, raw = load_testdata()
    bad = raw.astype(np.float32).copy()
    bad[7, 20, 30] = np.nan
    _, fbad = ob.reconstruct(p, otf, bad, 32, True)
    assert not np.isfinite(fbad[0].k0x)
    with _gpu(p, otf) as g:
        g.set_raw(bad)
        g.load_and_rescale(0)
        g.set_time(0)
        with pytest.raises(engine.SimrError, match="non-finite"):
            g.process_volume()
        g.set_raw(raw)
        g.load_and_rescale(0)
        g.set_time(0)
        g.process_volume()
        vol = g.result()
    ref, _ = ob.reconstruct(p, otf, raw, 32, True)
    assert rel_l2(vol, ref) < 1e-5


def test_pinned_host_buffers_give_the_same_result():
    """Stacks and results in page-locked host memory (simr_host_alloc) take the same path
    as pageable ones: bit-identical outputs."""
    log("pinned host buffers")
    p, otf, raw, _ = synthetic(64, 64, 8, nphases=5)
    raw = raw.astype(np.uint16)
    pin = engine.pinned_empty(raw.shape, np.uint16)
    pin[...] = raw
    with _gpu(p, otf) as g:
        g.set_raw(raw)
        g.load_and_rescale(0)
        g.process_volume()
        a = g.result()
        g.set_raw(pin)
        g.load_and_rescale(0)
        g.process_volume()
        out = engine.pinned_empty(a.shape, np.float32)
        b = g.result(out)
    assert b is out
    np.testing.assert_array_equal(a, b)
