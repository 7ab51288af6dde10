"""Pins of the makeotf oracle (oracle/makeotf_oracle.cpp, a restatement of
src/otf/radialft.cpp) against closed forms and invariants, on the CPU.

- the 1/n-normalised separation returns the phase mean as order 0;
- a widefield Gaussian bead gives the analytic Gaussian OTF (FFT, bead-centre
  phase ramp, radial binning and normalisation together);
- bead-size compensation divides by the textbook sphere transform
  3 (sin x - x cos x) / x^3, x = 2 pi R k;
- a 3-beam bead recovers the pattern phase of each order (combine_reim);
- fixorigin puts the low-kr row sums on the line fitted over [kr1, kr2];
- every output is Hermitian in kz with a real kz = 0 plane;
- cleanup zeroes everything beyond the lateral cut-off 2 NA / lambda.
"""
import numpy as np
import pytest

from cudasirecon_amd import synth
from cudasirecon_amd._abi import otf_default_params, set_otf_params
from oracle import binding as ob


def params(**kw):
    p = otf_default_params()
    set_otf_params(p, dr=0.08, dz=0.125, linespacing=0.2035, k0angle=-0.8, na=1.42, nimm=1.515, wavelength=515)
    return set_otf_params(p, **kw)


def gaussian_bead(nx, ny, nz, sx, sz, center):
    xc, yc, zc = center
    x = np.arange(nx)[None, None, :]
    y = np.arange(ny)[None, :, None]
    z = np.arange(nz)[:, None, None]
    return np.exp(-((x - xc) ** 2 + (y - yc) ** 2) / (2 * sx ** 2) - (z - zc) ** 2 / (2 * sz ** 2))


def radial_bins(nx, ny, dr):
    """Pixel -> bin map of the half spectrum as radialft builds it (radialft.cpp:893-914)."""
    nr = min(nx, ny)
    ky = np.fft.fftfreq(ny) * ny
    ky[ky > ny // 2] -= ny
    kx = np.arange(nx // 2 + 1)
    kyf = (ky * np.float32(1.0 / (ny * dr)))[:, None]
    kxf = (kx * np.float32(1.0 / (nx * dr)))[None, :]
    rd = np.sqrt(kxf ** 2 + kyf ** 2)
    inside = rd < 0.5 / dr
    b = np.rint(rd / (1.0 / (nr * dr))).astype(int)
    return b, inside, kxf, kyf


def test_order0_band_is_phase_mean():
    rng = np.random.default_rng(3)
    psf = rng.uniform(50, 500, size=(6 * 5, 32, 32)).astype(np.float32)
    p = params(nphases=5, napodize=0, background=20.0)
    _, _, b0 = ob.makeotf(p, psf, want_band0=True)
    ref = psf.reshape(6, 5, 32, 32).astype(np.float64).mean(axis=1) - 20.0
    assert np.abs(b0 - ref).max() < 2e-4 * np.abs(ref).max()


def test_widefield_gaussian_bead_gives_gaussian_otf():
    nx = ny = 64
    nz = 32
    sx, sz = 1.2, 1.6
    psf = gaussian_bead(nx, ny, nz, sx, sz, (31.3, 32.6, 15.4)).astype(np.float32) * 1000
    p = params(nphases=1, napodize=0, background=0.0, do_compen=0, ifixz=0)
    otf, info, _ = ob.makeotf(p, psf)
    assert abs(info.xc - 31.3) < 0.1 and abs(info.yc - 32.6) < 0.1 and abs(info.zc - 15.4) < 0.1
    b, inside, kxf, kyf = radial_bins(nx, ny, 0.08)
    fx, fy = kxf * 0.08, kyf * 0.08  # cycles per pixel
    g = np.exp(-2 * np.pi ** 2 * sx ** 2 * (fx ** 2 + fy ** 2))
    kz = np.fft.fftfreq(nz)
    gz = np.exp(-2 * np.pi ** 2 * sz ** 2 * kz ** 2)
    nkr = nx // 2 + 1
    expect = np.zeros((nkr, nz))
    for r in range(nkr):
        m = inside & (b == r)
        if m.any():
            expect[r] = g[m].mean() * gz
    got = otf[0]
    assert np.abs(np.abs(got) - expect).max() < 3e-3
    # the phase ramp of the sub-pixel bead centre is removed at low frequencies
    low = expect > 0.2
    assert np.abs(np.angle(got[low])).max() < 0.1


def test_bead_compensation_divides_by_sphere_transform():
    nx = ny = 64
    nz = 32
    bead = gaussian_bead(nx, ny, nz, 1.0, 1.5, (32.0, 32.0, 16.0)).astype(np.float32) * 1000
    # three identical phases: order 0 is the widefield bead, k0 of order 0 is 0
    psf = np.repeat(bead[:, None], 3, axis=1).reshape(nz * 3, ny, nx)
    p0 = params(nphases=3, napodize=0, background=0.0, do_compen=0, ifixz=0, bead_diameter=0.2)
    p1 = params(nphases=3, napodize=0, background=0.0, do_compen=1, ifixz=0, bead_diameter=0.2)
    off, _, _ = ob.makeotf(p0, psf)
    on, _, _ = ob.makeotf(p1, psf)
    R = 0.1
    b, inside, kxf, kyf = radial_bins(nx, ny, 0.08)
    k = np.sqrt(kxf ** 2 + kyf ** 2)
    x = 2 * np.pi * R * np.maximum(k, 1e-12)
    sphere = 3 * (np.sin(x) - x * np.cos(x)) / x ** 3
    for r in range(1, 12):
        m = inside & (b == r)
        expect = np.mean(1.0 / sphere[m])
        got = on[0, r, 0].real / off[0, r, 0].real
        assert abs(got / expect - 1) < 3e-3, (r, got, expect)


def test_single_phase_skips_bead_compensation():
    # with nphases = 1 the order-0 k0 is 0/(norders-1) = NaN, so sphereFFT
    # takes its k == 0 branch everywhere and the ratio is 1 (radialft.cpp:809-832)
    psf = gaussian_bead(32, 32, 16, 1.0, 1.5, (16.0, 16.0, 8.0)).astype(np.float32) * 1000
    off, _, _ = ob.makeotf(params(nphases=1, do_compen=0), psf)
    on, _, _ = ob.makeotf(params(nphases=1, do_compen=1), psf)
    np.testing.assert_array_equal(on, off)


@pytest.mark.parametrize("phase0", [0.3, 0.9, -0.4])
def test_side_band_phase_is_recovered(phase0):
    psf = synth.make_psf_stack(64, 64, 24, 5, seed=1, phase0=phase0, noise=False)
    p = params(nphases=5, background=100.0)
    otf, info, _ = ob.makeotf(p, psf)
    assert abs(info.phi[1] - phase0) < 0.02
    assert abs(info.phi[2] - 2 * phase0) < 0.02
    # the combined side bands are then (nearly) real at low kr, kz = 0
    assert abs(np.angle(otf[1, 2, 0])) < 0.05 and abs(np.angle(otf[2, 2, 0])) < 0.05


def test_fixorigin_puts_low_kr_row_sums_on_the_fitted_line():
    psf = synth.make_psf_stack(64, 64, 24, 5, seed=2)
    p = params(nphases=5, interpkr=[3, 20], ifixz=0)
    otf, _, _ = ob.makeotf(p, psf)
    sums = otf[0].real.astype(np.float64).sum(axis=1)
    sums[0] = otf[0, 0, 0].real
    i = np.arange(3, 21)
    slope, icpt = np.polyfit(i - 0.5 * (3 + 20), sums[3:21], 1)
    for j in range(3):
        assert abs(sums[j] - (icpt + slope * (j - 11.5))) < 1e-5 * np.abs(sums).max()


@pytest.mark.parametrize("nphases,nz", [(5, 24), (3, 17), (1, 16)])
def test_output_is_hermitian_in_kz(nphases, nz):
    psf = synth.make_psf_stack(48, 48, nz, nphases, seed=4)
    p = params(nphases=nphases)
    otf, _, _ = ob.makeotf(p, psf)
    for band in otf:
        assert np.abs(band[:, 0].imag).max() == 0
        conj = np.conj(band[:, (-np.arange(nz)) % nz])
        assert np.abs(band - conj).max() <= 1e-6 * np.abs(band).max()


def test_cleanup_zeroes_beyond_the_lateral_cutoff():
    psf = synth.make_psf_stack(64, 64, 24, 5, seed=5)
    p = params(nphases=5, leavekz=[7, 11, 3])
    otf, info, _ = ob.makeotf(p, psf)
    krmax = 2 * 1.42 / 0.515
    kr = np.arange(otf.shape[1]) * np.float32(1.0 / (64 * 0.08))
    assert np.all(otf[:, kr > krmax * (1 + 1e-6), :] == 0)
    assert np.any(otf[:, kr < krmax * 0.5, :] != 0)
