"""Pins of the CPU oracle (oracle/) against closed forms, numpy and brute force.

The reference holds no expected outputs (its test_data outputs are gitignored,
.gitignore:5-7 of the reference), so every oracle function is pinned here by an
independent check chosen so that a dropped term, a sign or an index slip fails.
"""
import math

import numpy as np
import pytest

from cases import rel_l2, synthetic, load_testdata
from cudasirecon_amd._abi import default_params, set_params
from oracle import binding as ob


# ------------------------------------------------------------------ FFT
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 15, 16, 21, 24, 25, 27, 40, 48, 49, 64, 100, 131,
                               262, 384, 640, 1024])
def test_fft_matches_numpy_1d(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for inv in (False, True):
        ref = np.fft.ifft(x) * n if inv else np.fft.fft(x)
        y = ob.fft3d(x.astype(np.complex128), inverse=inv)
        assert np.abs(y - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
        yf = ob.fft3d(x.astype(np.complex64), inverse=inv)
        assert rel_l2(yf, ref) < 2e-6


def test_fft_brute_force_dft():
    for n in (3, 5, 7, 9, 12, 15):
        rng = np.random.default_rng(100 + n)
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        k = np.arange(n)
        W = np.exp(-2j * np.pi * np.outer(k, k) / n)
        assert np.allclose(ob.fft3d(x), W @ x, atol=1e-12)


def test_fft_3d_and_parseval():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((6, 10, 12)) + 1j * rng.standard_normal((6, 10, 12))
    y = ob.fft3d(x)
    assert np.allclose(y, np.fft.fftn(x), atol=1e-11)
    assert math.isclose(np.sum(np.abs(y) ** 2), x.size * np.sum(np.abs(x) ** 2), rel_tol=1e-12)


# ------------------------------------------------------------------ scalar helpers
def test_fitparabola_exact_on_quadratic():
    # y = -(x - 0.3)^2 + 5 sampled at -1, 0, 1 -> peak at +0.3
    f = lambda x: -(x - 0.3) ** 2 + 5  # noqa: E731
    assert abs(ob.fitparabola(f(-1), f(0), f(1)) - 0.3) < 1e-6
    # flat -> 0; peak outside +-1.5 -> 0 (gpuFunctionsImpl.cu:485-493)
    assert ob.fitparabola(1.0, 1.0, 1.0) == 0.0
    g = lambda x: -(x - 3.0) ** 2  # noqa: E731
    assert ob.fitparabola(g(-1), g(0), g(1)) == 0.0


def test_fitxyparabola_exact():
    f = lambda x: -2.5 * (x - 1.234) ** 2 + 7  # noqa: E731
    xs = [0.9, 1.3, 1.45]
    assert abs(ob.fitxyparabola(xs[0], f(xs[0]), xs[1], f(xs[1]), xs[2], f(xs[2])) - 1.234) < 1e-9
    assert ob.fitxyparabola(1, 2, 1, 3, 2, 4) == 0.0  # two equal x -> 0


def test_findpeak_first_max_and_wrap():
    a = np.zeros((8, 10), np.float32)
    a[0, 0] = 5.0
    a[0, 9] = 3.0  # left neighbour through the wrap
    a[0, 1] = 4.0
    a[7, 0] = 1.0
    a[1, 0] = 2.0
    px, py = ob.findpeak(a)
    # x: fitparabola(3, 5, 4) = -0.5*(4-3)/(4+3-10) ; y: fitparabola(1, 5, 2)
    assert abs(px - (0.5 * 1 / 3)) < 1e-6
    assert abs(py - (0.5 * 1 / 7)) < 1e-6
    b = np.ones((4, 4), np.float32)
    b[2, 1] = b[3, 3] = 9.0  # tie: the first in row-major order wins
    px, py = ob.findpeak(b)
    assert round(px) == 1 and round(py) == 2


# ------------------------------------------------------------------ separation matrix
@pytest.mark.parametrize("nphases", [3, 5])
def test_makematrix_inverts_forward_model(nphases):
    p, tab, raw, _ = synthetic(32, 32, 4, nphases=nphases)
    o = ob.Oracle(p, tab, 64)
    M = o.sepmatrix()
    norders = nphases // 2 + 1
    phi = 2 * np.pi * np.arange(nphases) / nphases
    # forward model img_j = a0 + sum_o c_o cos(o phi_j) + s_o sin(o phi_j)
    F = [np.ones(nphases)]
    for order in range(1, norders):
        F += [np.cos(order * phi), np.sin(order * phi)]
    F = np.array(F).T
    expect = np.diag([nphases] + [nphases / 2.0] * (2 * norders - 2))
    assert np.allclose(M @ F, expect, atol=1e-12)


# ------------------------------------------------------------------ preprocessing
def test_rescale_equalises_plane_sums():
    p, tab, raw, _ = synthetic(32, 32, 4)
    o = ob.Oracle(p, tab, 64)
    o.load(raw)
    o.rescale(0)
    nx, ny, nz = p.nx, p.ny, p.nz
    planes = []
    for d in range(p.ndirs):
        for ph in range(p.nphases):
            b = o.band(d, ph, double=True).view(np.float64).reshape(nz, ny, (nx // 2 + 1) * 2)[:, :, :nx]
            planes.append(b.sum(axis=(1, 2)))
    ref = planes[0]
    for s in planes:
        assert np.allclose(s, ref, rtol=1e-12)


def test_inscale_formula():
    p, tab, raw, _ = synthetic(32, 32, 4)
    o = ob.Oracle(p, tab, 64)
    assert math.isclose(o.inscale(), 1.0 / (32 * 32 * 4 * 2 * 2 * 1 * 3), rel_tol=1e-15)


# ------------------------------------------------------------------ Wiener filter special case
def _linear_ra_otf(nxotf, nzotf, slope):
    """OTF linear in kr and constant in kz: bilinear interpolation reproduces it exactly."""
    kr = np.arange(nxotf, dtype=np.float64)
    v = np.clip(1.0 - slope * kr, 0, None)
    return np.repeat(v[:, None], nzotf, axis=1).astype(np.complex64)[None]


def test_filter_reduces_to_textbook_wiener():
    # ndirs = nphases = 1, no suppression, no apodisation: scale = conj(H)/(|H|^2 + w^2)
    nx = ny = 32
    nz = 4
    p = default_params()
    nzotf, nxotf, dkr = 8, 65, 0.2
    set_params(p, nx=nx, ny=ny, nz=nz, ndirs=1, nphases=1, dxy=0.1, dz=0.2, wave=500, na=1.2, nimm=1.33,
               bRadAvgOTF=1, nzotf=nzotf, nxotf=nxotf, nyotf=1, dkrotf=dkr, dkzotf=1.0 / (0.2 * nz),
               bSuppress_singularities=0, apodizeoutput=0, wiener=0.05, napodize=0, k0angles=[0.0])
    otf = _linear_ra_otf(nxotf, nzotf, 0.01)
    o = ob.Oracle(p, otf, 64)
    rng = np.random.default_rng(0)
    band = (rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1)))
    o.set_band(0, 0, band.astype(np.complex64))
    o.filter_dir(0)
    got = o.band(0, 0, double=True)
    # expected, element by element (filterbands_kernel1 with a single band)
    dkx = 1.0 / (nx * 0.1)
    rcut = min(2 * 1.2 / 0.5, 1 / (2 * 0.1))
    zd = o.zdist()[0]
    exp = np.zeros_like(band)
    for iz in range(nz):
        z0 = iz if iz <= zd else iz - nz
        if not (iz <= zd or iz >= nz - zd):
            continue
        for iy in range(ny):
            y1 = iy if iy < ny // 2 else iy - ny
            for ix in range(nx // 2 + 1):
                kr = math.hypot(ix * dkx, y1 * dkx)
                if kr > rcut:
                    continue
                H = max(1.0 - 0.01 * kr / dkr, 0.0)
                exp[iz, iy, ix] = band.astype(np.complex64)[iz, iy, ix] * H / (H * H + 0.05 ** 2)
    assert rel_l2(got, exp) < 1e-6


def _multilinear_full_otf(nxotf, nyotf, nzotf, dkr, dkz):
    """A full OTF table [1][kz][ky][kx] (kx >= 0; ky, kz in FFT order) of the complex
    function H = (1 - 0.01 kx/dkr)(1 + 0.004 ky/dkr)(1 + 0.05i kz/dkz), multilinear in
    (kx, ky, kz), so trilinear interpolation reproduces it exactly inside the table."""
    kx = np.arange(nxotf)
    ky = np.fft.fftfreq(nyotf, 1.0 / nyotf)
    kz = np.fft.fftfreq(nzotf, 1.0 / nzotf)
    h = ((1 - 0.01 * kx)[None, None, :] * (1 + 0.004 * ky)[None, :, None] * (1 + 0.05j * kz)[:, None, None])
    return h.astype(np.complex64)[None], lambda x, y, z: (1 - 0.01 * x / dkr) * (1 + 0.004 * y / dkr) * (1 + 0.05j * z / dkz)


@pytest.mark.parametrize("nz", [4, 1])
def test_filter_full_otf_reduces_to_textbook_wiener(nz):
    """A full (not rotationally averaged) OTF: the same single-band Wiener case, its OTF
    values read by trilinear interpolation at off-grid (kx, ky, kz) of a multilinear table."""
    nx = ny = 32
    dkr = 0.23
    nxotf, nyotf = 40, 96
    nzotf = 8 if nz > 1 else 1
    dkzotf = 1.0 / (0.2 * 6.5) if nz > 1 else 1.0
    p = default_params()
    set_params(p, nx=nx, ny=ny, nz=nz, ndirs=1, nphases=1, dxy=0.1, dz=0.2, wave=500, na=1.2, nimm=1.33,
               bRadAvgOTF=0, nzotf=nzotf, nxotf=nxotf, nyotf=nyotf, dkrotf=dkr, dkzotf=dkzotf,
               bSuppress_singularities=0, apodizeoutput=0, wiener=0.05, napodize=0, k0angles=[0.0])
    otf, H = _multilinear_full_otf(nxotf, nyotf, nzotf, dkr, dkzotf)
    o = ob.Oracle(p, otf, 64)
    rng = np.random.default_rng(2)
    band = (rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1)))
    o.set_band(0, 0, band.astype(np.complex64))
    o.filter_dir(0)
    got = o.band(0, 0, double=True)
    dkx = 1.0 / (nx * 0.1)
    dkz = 1.0 / (nz * 0.2)
    rcut = min(2 * 1.2 / 0.5, 1 / (2 * 0.1))
    zd = o.zdist()[0]
    exp = np.zeros_like(band)
    for iz in range(nz):
        z0 = iz if iz <= zd else iz - nz
        if not (iz <= zd or iz >= nz - zd):
            continue
        kz = z0 * dkz if nz > 1 else 0.0
        for iy in range(ny):
            y1 = iy if iy < ny // 2 else iy - ny
            for ix in range(nx // 2 + 1):
                if math.hypot(ix * dkx, y1 * dkx) > rcut:
                    continue
                h = H(ix * dkx, y1 * dkx, kz)
                exp[iz, iy, ix] = band.astype(np.complex64)[iz, iy, ix] * np.conj(h) / (abs(h) ** 2 + 0.05 ** 2)
    assert rel_l2(got, exp) < 1e-6


@pytest.mark.parametrize("nz,nphases", [(8, 5), (1, 3)])
def test_full_otf_sampled_from_ra_reproduces_ra_reconstruction(nz, nphases):
    """A full OTF table that samples an RA OTF on the (kx, ky, kz) grid: the reconstruction
    (fit, filter, assembly, every lookup through the Hermitian mirror for kx < 0 and the
    ky / kz wraps) equals the RA one up to the interpolation difference (trilinear on the
    grid versus linear in kr), and the fitted k0 agree.  In 3D the RA table carries the
    axial phase of a PSF 0.37 planes off centre, so it is complex and Hermitian, and a
    mirror that lost its conjugation would differ at every kx < 0."""
    from cases import full_otf_from_ra

    p, tab, raw, _ = synthetic(48, 48, nz, nphases=nphases)
    tab = np.asarray(tab)
    if nz > 1:
        nzotf = tab.shape[2]
        kz = np.fft.fftfreq(nzotf, 1.0 / nzotf)
        tab = (tab * np.exp(-2j * np.pi * kz * 0.37 / nzotf)[None, None, :]).astype(np.complex64)
    q, full = full_otf_from_ra(p, tab)
    outs, fits = [], []
    for par, t in ((p, tab), (q, full)):
        o = ob.Oracle(par, t, 64)
        o.load(raw.astype(np.float32))
        o.rescale(0)
        o.set_time(0)
        o.process_volume()
        outs.append(o.result(double=True))
        fits.append([o.fit(d) for d in range(p.ndirs)])
    assert rel_l2(outs[1], outs[0]) < 3e-3
    dk = 1.0 / (p.nx * p.dxy)
    for a, b in zip(*fits):
        assert math.hypot(a.k0x - b.k0x, a.k0y - b.k0y) / dk < 1e-3


# ------------------------------------------------------------------ assembly special case
def _moved_grid(full, nx, ny, nz, Z):
    """move_kernel's scatter (gpuFunctionsImpl.cu:1853-1913) of a full spectrum into the zoomed grid."""
    big = np.zeros((nz, Z * ny, Z * nx), np.complex128)
    for zc in range(nz):
        z = zc - (nz // 2 - 1)
        for yc in range(ny):
            y = yc - (ny // 2 - 1)
            for xc in range(nx):
                x = xc - (nx // 2 - 1)
                big[z % nz, y % (Z * ny), x % (Z * nx)] = full[z % nz, y % ny, x % nx]
    return big


def test_assembly_order0_is_fourier_interpolation():
    """ndirs = nphases = 1 with a unit OTF, w = 0 and no apodisation: the filter only applies
    its support masks, and the output is the zero-padded (Fourier) interpolation of the
    masked spectrum, Re IFFT of the zoomed grid."""
    nx = ny = 16
    nz = 6
    dxy, dz = 0.1, 0.2
    p = default_params()
    set_params(p, nx=nx, ny=ny, nz=nz, ndirs=1, nphases=1, dxy=dxy, dz=dz, wave=500, na=1.2, nimm=1.33,
               bRadAvgOTF=1, nzotf=8, nxotf=65, nyotf=1, dkrotf=0.2, dkzotf=1.0 / (dz * nz), napodize=0,
               k0angles=[0.0], wiener=0.0, apodizeoutput=0, bSuppress_singularities=0)
    otf = _linear_ra_otf(65, 8, 0.0)
    o = ob.Oracle(p, otf, 64)
    rng = np.random.default_rng(1)
    real = rng.standard_normal((nz, ny, nx))
    full = np.fft.fftn(real)
    o.set_band(0, 0, full[:, :, : nx // 2 + 1].astype(np.complex64))
    o.set_fit(0, o.fit(0))
    o.filter_assemble()
    got = o.result(double=True)
    zd = o.zdist()[0]
    dk = 1.0 / (nx * dxy)
    rcut = min(2 * 1.2 / 0.5, 1 / (2 * dxy))
    ky = np.fft.fftfreq(ny, 1.0 / ny)
    kx = np.fft.fftfreq(nx, 1.0 / nx)
    ky = np.where(ky >= ny // 2, ky - ny, ky)  # y1 in [-ny/2, ny/2-1]
    kx = np.abs(kx)
    mask_r = np.hypot(kx[None, :] * dk, ky[:, None] * dk) <= rcut
    iz = np.arange(nz)
    mask_z = (iz <= zd) | (iz >= nz - zd)
    masked = full.astype(np.complex64).astype(np.complex128) * mask_r[None] * mask_z[:, None, None]
    big = _moved_grid(masked, nx, ny, nz, 2)
    expect = np.real(np.fft.ifftn(big)) * big.size
    assert got.shape == expect.shape
    assert rel_l2(got, expect) < 1e-6


# ------------------------------------------------------------------ end to end
def test_testdata_k0_near_pattern_geometry():
    """raw.dv: k0 lands next to the configured angles and line spacing (~12.6 px at 64^2)."""
    p, otf, raw = load_testdata()
    vol, fits = ob.reconstruct(p, otf, raw, 64)
    dk = 1.0 / (p.nx * p.dxy)
    for f in fits:
        mag = math.hypot(f.k0x, f.k0y) / dk
        assert abs(mag - (1 / 0.2035 / 2) / dk) < 0.1
        assert math.hypot(f.k0x - f.k0guess_x, f.k0y - f.k0guess_y) / dk < 0.05
        assert 0.3 < abs(complex(f.amp_re[1], f.amp_im[1])) < 1.5
        assert f.corr[2] > 0.9
    assert np.isfinite(vol).all() and vol.shape == (9, 128, 128)


@pytest.mark.parametrize("nz,nphases", [(8, 5), (1, 3)])
def test_synthetic_fit_recovers_ground_truth(nz, nphases):
    p, tab, raw, truth = synthetic(64, 64, nz, nphases=nphases, photons=None)
    vol, fits = ob.reconstruct(p, tab, raw.astype(np.float32), 64)
    dk = 1.0 / (p.nx * p.dxy)
    for d, f in enumerate(fits):
        err = math.hypot(f.k0x - truth["k0"][d][0], f.k0y - truth["k0"][d][1]) / dk
        assert err < (0.02 if nz > 1 else 0.05), (d, err)
        for o, m in enumerate(truth["modamps"], start=1):
            assert abs(abs(complex(f.amp_re[o], f.amp_im[o])) - m) < 0.03 * max(1, m / 0.2)


def test_fp32_oracle_close_to_fp64_oracle():
    p, otf, raw = load_testdata()
    v32, f32 = ob.reconstruct(p, otf, raw, 32)
    v64, f64 = ob.reconstruct(p, otf, raw, 64)
    assert rel_l2(v32, v64) < 2e-5
    dk = 1.0 / (p.nx * p.dxy)
    for a, b in zip(f32, f64):
        assert math.hypot(a.k0x - b.k0x, a.k0y - b.k0y) / dk < 1e-3


# ------------------------------------------------------------------ Bessel-sheet branches
def test_bessel_2d_is_the_plain_path_with_otfcutoff_001():
    """In 2D every Bessel branch but the overlap OTF cutoff collapses: fitorder2 is
    1 either way, order0_2_factor is 1, every axial cutoff clamps to 0 and the
    apodization takes its 2D form (gpuFunctionsImpl.cu:316-323, 505, 548, 554,
    1286-1287, 1566-1586).  So --bessel in 2D equals the plain path run with
    otfcutoff 0.01, and differs from the plain path with the default 0.006."""
    p, tab, raw, _ = synthetic(64, 64, 1, nphases=3, photons=None)
    plain = p.copy()
    plain.otfcutoff = 0.01
    bes = p.copy()
    bes.bBessel, bes.BesselNA, bes.BesselLambdaEx = 1, 1.3, 0.488  # kzExMax 5.33 > |k0| 4.91
    va, fa = ob.reconstruct(plain, tab, raw.astype(np.float32), 32)
    vb, fb = ob.reconstruct(bes, tab, raw.astype(np.float32), 32)
    np.testing.assert_array_equal(va, vb)
    for a, b in zip(fa, fb):
        assert (a.k0x, a.k0y, list(a.amp_re)) == (b.k0x, b.k0y, list(b.amp_re))
    v0, _ = ob.reconstruct(p, tab, raw.astype(np.float32), 32)
    assert rel_l2(v0, vb) > 0


def _bessel3d(**kw):
    # |k0| = 1/0.2035/2 = 2.457/um against kzExMax = 2*0.65/0.488 = 2.664/um;
    # dz 0.1 keeps the axial cutoffs (15, 14, 10 for orders 0, 1, 2) below nz/2
    return synthetic(64, 64, 32, nphases=5, dz=0.1, bBessel=1, BesselNA=0.65, BesselLambdaEx=0.488, **kw)


def test_bessel_3d_fit_on_order_1_recovers_ground_truth():
    """--bessel fits k0 on the (0, 1) overlap and divides the cross-correlation
    peak by 1, not 2 (findk0, gpuFunctionsImpl.cu:316-372): the fitted k0 must
    land on the simulated first-order frequency and the modamps on the truth."""
    p, tab, raw, truth = _bessel3d(photons=None)
    vol, fits = ob.reconstruct(p, tab, raw.astype(np.float32), 64)
    dk = 1.0 / (p.nx * p.dxy)
    for d, f in enumerate(fits):
        err = math.hypot(f.k0x - truth["k0"][d][0], f.k0y - truth["k0"][d][1]) / dk
        # an order-1 fit has half the order-2 fit's lever: 0.05 px (halving k0 would miss by ~6 px)
        assert err < 0.05, (d, err)
        for o, m in enumerate(truth["modamps"], start=1):
            assert abs(abs(complex(f.amp_re[o], f.amp_im[o])) - m) < 0.03 * max(1, m / 0.2), (d, o)
    assert np.isfinite(vol).all()


def test_bessel_refuses_a_pattern_beyond_the_excitation_kz_extent():
    """cos(halfangle) = |k0| order / (norders-1) / kzExMax above 1 has no angle
    (the reference takes acosf of it and truncates a NaN): a clear error."""
    p, tab, raw, _ = _bessel3d(photons=None)
    q = p.copy()
    q.BesselNA = 0.3  # kzExMax 1.23 < |k0| 2.457
    with pytest.raises(Exception, match="Bessel"):
        ob.reconstruct(q, tab, raw.astype(np.float32), 32)


def test_bessel_filter_axial_cutoffs_at_known_half_angles():
    """The filter's Bessel axial cutoffs (gpuFunctionsImpl.cu:1260-1271) where the
    half angle is exact: with besselExWave 2 and besselNA = |k0|, kzExMax = |k0|,
    so order 2 of 3 sits at cos(halfangle) = 1 and keeps no excitation kz extent
    (its cutoff is the plain path's detection-only zdistcutoff[0]), order 1 sits
    at 60 degrees (sin = sqrt(3)/2), and order 0 spans kzExMax plus the detection
    extent, rounded to nearest."""
    p, tab, _, _ = synthetic(64, 64, 32, nphases=5, dz=0.1, photons=None)
    k = float(np.float32(2.4))

    def cutoffs(q):
        o = ob.Oracle(q, tab, 32)
        for d in range(q.ndirs):
            f = o.fit(d)
            f.k0x, f.k0y = k, 0.0
            for i in range(3):
                f.amp_re[i], f.amp_im[i] = 1.0, 0.0
            o.set_fit(d, f)
        o.filter_dir(0)
        return o.zdist()

    plain = cutoffs(p)
    q = p.copy()
    q.bBessel, q.BesselNA, q.BesselLambdaEx = 1, k, 2.0
    bes = cutoffs(q)
    lam = 515 / 1.515 / 1000.0
    det = (1 - math.cos(math.asin(1.42 / 1.515))) / lam  # detection OTF's axial extent, 1/um
    dkz = 1 / (32 * 0.1)
    assert bes[2] == plain[0] == math.ceil(det / dkz)
    assert bes[1] == math.ceil((k * math.sqrt(3) / 2 + det) / dkz)
    assert bes[0] == round((k + det) / dkz)
    assert len(set(bes)) == 3  # three distinct cutoffs: an order mix-up shows


# ------------------------------------------------------------------ deskew / cropXY resampling
def _loaded(q, tab, raw):
    """Flat-fielded, resampled sections the oracle loads: [dir][phase] -> (nz, ny, nx)."""
    o = ob.Oracle(q, tab, 64)
    o.load(raw)
    secs = [[o.band(d, ph, double=True).view(np.float64).reshape(q.nz, q.ny, -1)[:, :, :q.nx]
             for ph in range(q.nphases)] for d in range(q.ndirs)]
    return secs, o.inscale()


def _flat(r, bg, inscale):
    # load_and_flatfield in float: ((raw - bg) - 0) * slope * inscale, slope 1
    return ((r.astype(np.float32) - np.float32(bg)) - np.float32(0)) * np.float32(1) * np.float32(inscale)


def test_crop_keeps_the_left_columns():
    """cropXY without deskew (deskewOneSection with factor 0, cudaSirecon.cpp:1743-1745):
    output column x is raw column x of the flat-fielded section."""
    p, tab, _, _ = synthetic(64, 64, 8, nphases=5, photons=None)
    q = p.copy()
    q.nx_raw = 96
    rng = np.random.default_rng(3)
    raw = rng.integers(80, 4000, size=(3 * 5 * 8, 64, 96)).astype(np.float32)
    secs, inscale = _loaded(q, tab, raw)
    for d in range(3):
        for z in range(8):
            for ph in range(5):
                exp = _flat(raw[(d * 8 + z) * 5 + ph], p.constbkgd, inscale)[:, :64]
                np.testing.assert_array_equal(secs[d][ph][z], exp.astype(np.float64))


def test_deskew_by_whole_pixels_is_an_exact_shift():
    """deskew factor 1 (cudaSirecon.cpp:1746-1757): section z is read at
    x - nx/2 + shift - (z - nz/2) + nx_raw/2, an integer, so the interpolation
    returns the raw value there exactly, and 0 outside the section."""
    p, tab, _, _ = synthetic(64, 64, 8, nphases=5, photons=None)
    q = p.copy()
    q.nx_raw, q.deskew_factor, q.deskew_shift = 64, 1.0, 3
    rng = np.random.default_rng(4)
    raw = rng.integers(80, 4000, size=(3 * 5 * 8, 64, 64)).astype(np.float32)
    secs, inscale = _loaded(q, tab, raw)
    for d in range(3):
        for z in range(8):
            for ph in range(5):
                f = _flat(raw[(d * 8 + z) * 5 + ph], p.constbkgd, inscale)
                exp = np.zeros((64, 64), np.float32)
                for x in range(64):
                    xin = x - 32 + 3 - (z - 4) + 32
                    if 0 <= xin < 64:
                        exp[:, x] = f[:, xin]
                np.testing.assert_array_equal(secs[d][ph][z], exp.astype(np.float64))


def test_deskew_interpolation_reproduces_a_linear_ramp():
    """deskew factor 0.5 puts odd sections at half pixels; linear interpolation of
    a ramp along x returns the ramp at the fractional position (away from the
    last raw column, whose right neighbour is the next row's first pixel)."""
    p, tab, _, _ = synthetic(64, 64, 8, nphases=5, photons=None)
    q = p.copy()
    q.nx_raw, q.deskew_factor, q.deskew_shift = 64, 0.5, 0
    xs = np.arange(64, dtype=np.float64)
    row = 200.0 + 30.0 * xs
    raw = np.broadcast_to(row.astype(np.float32), (3 * 5 * 8, 64, 64)).copy()
    secs, inscale = _loaded(q, tab, raw)
    for z in range(8):
        xin = xs - 0.5 * (z - 4)
        ok = (xin >= 0) & (xin < 63)
        exp = (200.0 + 30.0 * xin - p.constbkgd) * inscale
        got = secs[1][2][z]
        assert np.allclose(got[:, ok], exp[ok], rtol=2e-6, atol=0), z
        assert (got[:, xin < 0] == 0).all()


# ------------------------------------------------------------------ flat field
@pytest.mark.parametrize("prec", [32, 64])
def test_flatfield_maps_undo_a_known_camera_response(prec):
    """--usecorr (bgAndSlope, cudaSirecon.cpp:139-192; load_and_flatfield
    1715-1731): each raw pixel becomes (raw - background) * slope.  A stack
    recorded through a camera with offset b(x, y) and gain 1/s(x, y), i.e.
    raw = b + clean / s, reconstructs with the maps (b, s) as the clean stack
    does with background 0 and unit slope."""
    p, otf, raw, _ = synthetic(64, 64, 8, seed=21)
    clean = raw.astype(np.float64) - p.constbkgd
    rng = np.random.default_rng(4)
    b = rng.uniform(80.0, 120.0, (p.ny, p.nx))
    s = rng.uniform(0.8, 1.25, (p.ny, p.nx))
    recorded = (b + clean / s).astype(np.float32)
    q = p.copy()
    q.constbkgd = 0.0
    o_ref = ob.Oracle(q, otf, prec, True)
    o_ref.load(clean.astype(np.float32))
    o_ref.rescale(0)
    o_ref.set_time(0)
    o_ref.process_volume()
    o = ob.Oracle(q, otf, prec, True)
    o.set_flatfield(b.astype(np.float32), s.astype(np.float32))
    o.load(recorded)
    o.rescale(0)
    o.set_time(0)
    o.process_volume()
    for d in range(p.ndirs):
        f, g = o.fit(d), o_ref.fit(d)
        assert abs(f.k0x - g.k0x) < 1e-4 and abs(f.k0y - g.k0y) < 1e-4
    assert rel_l2(o.result(double=prec == 64), o_ref.result(double=prec == 64)) < 2e-6
    # a wrong map (background left out) does not pass
    o_bad = ob.Oracle(q, otf, prec, True)
    o_bad.set_flatfield(np.zeros_like(b, np.float32), s.astype(np.float32))
    o_bad.load(recorded)
    o_bad.rescale(0)
    o_bad.set_time(0)
    o_bad.process_volume()
    assert rel_l2(o_bad.result(), o_ref.result()) > 1e-3
