"""Pins of the oracle's sideband assembly and generalised Wiener filter by what
the mathematics of SIM fixes, independent of the reference's code:

  * assembly (assemblerealspacebands, gpuFunctionsImpl.cu:1741-1954) is the
    shift theorem: a band holding one spectral delta assembles into the
    analytic cosine at that frequency plus order * k0;
  * the filter's denominator (filterbands_kernel1, gpuFunctionsImpl.cu:
    1443-1656) with a unit OTF is the number of bands whose support covers
    the absolute frequency, each weighted by its |modamp|^2, plus wiener^2;
  * singularity suppression, order-0 damping and the output apodization
    (dev_suppress / dev_order0damping 1717-1739, 1579-1597) at points where
    each factor has a closed value;
  * a noiseless stack simulated with a flat OTF, known k0 and modamps
    reconstructs to the sample's spectrum inside the extended support,
    weighted by the Wiener factor, and to zero outside it.

Each check runs on the FP32 and the FP64 oracle.  The cases are chosen so
that a dropped term, a sign of the sideband phasor, a missing conjugation or
a transposed index fails.
"""
import math

import numpy as np
import pytest

from cudasirecon_amd._abi import SimrFit, default_params, set_params
from oracle import binding as ob

DXY = float(np.float32(0.1))  # um, as the parameter block stores it


def pin_params(nx, ny, nz, ndirs, nphases, dz=0.2, **kw):
    """A 2D/3D geometry with a flat-OTF table and every optional factor off
    (no output apodization, suppression, order-0 damping, edge blending or
    bleach correction); wave 500 nm and NA 0.5586 put the OTF cut-off at
    2 NA / lambda = 2.2344 1/um (14.3 pixels of a 64-pixel grid)."""
    p = default_params()
    nzotf = 1 if nz == 1 else 16
    set_params(p, nx=nx, ny=ny, nz=nz, ndirs=ndirs, nphases=nphases, dxy=DXY, dz=dz if nz > 1 else 0.0, wave=500,
               na=0.5586, nimm=1.515, bRadAvgOTF=1, nxotf=64, nyotf=1, nzotf=nzotf, dkrotf=0.1,
               dkzotf=1.0 / (nzotf * 0.2), apodizeoutput=0, bSuppress_singularities=0, bDampenOrder0=0, napodize=0,
               do_rescale=0, constbkgd=0.0, wiener=0.1)
    set_params(p, **kw)
    return p


def unit_otf(p):
    return np.ones((p.nphases // 2 + 1, p.nxotf, p.nzotf), np.complex64)


def rcut_of(p):
    """The OTF support radius 2 NA / lambda in 1/um, as the reference rounds it (float)."""
    return float(np.float32(np.float32(p.na) * np.float32(2.0)) / (p.wave / 1000.0))


def set_fit(o, d, k0, amps):
    f = SimrFit()
    f.k0x, f.k0y = k0
    for i, a in enumerate(amps):
        f.amp_re[i], f.amp_im[i] = a.real, a.imag
    o.set_fit(d, f)


def _tol(prec):
    return 1e-5 if prec == 32 else 1e-10


# ----------------------------------------------------------------- assembly
@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("nz,zz", [(1, 1), (8, 1), (6, 2)])
def test_assembly_is_the_shift_theorem(prec, nz, zz):
    """Band 0 holding A at k assembles to 2|A| cos(2 pi k.r + arg A); the
    order-o side band holding only its positive half (re = A, im = -i A, so
    re + i im = 2A at k and 0 at -k) assembles to 4|A| cos(2 pi (k + o k0).r
    + arg A), with r measured from the zoomed grid's centre for the k0 term."""
    nx, ny = 32, 24
    ndirs = 2
    p = pin_params(nx, ny, nz, ndirs, 5, z_zoom=zz)
    o = ob.Oracle(p, unit_otf(p), prec)
    nxh = nx // 2 + 1
    k0 = [(np.float32(1.37), np.float32(-0.52)), (np.float32(-0.61), np.float32(1.13))]
    kz = (lambda v: v) if nz > 1 else (lambda v: 0)
    deltas = [  # (direction, order, value, (kx, ky, kz) in grid units)
        (0, 0, 0.9 * np.exp(0.3j), (3, -2, kz(1))),
        (0, 1, 0.5 * np.exp(-1.1j), (5, 4, kz(-2))),
        (0, 2, 0.25 * np.exp(2.0j), (2, -7, kz(0))),
        (1, 1, 0.4 * np.exp(0.7j), (7, 1, kz(2))),
        (1, 2, 0.3 * np.exp(-2.5j), (1, 9, kz(-1))),
        (1, 0, 0.6 * np.exp(-0.4j), (6, 3, kz(0))),
    ]
    bands = {(d, b): np.zeros((nz, ny, nxh), np.complex64) for d in range(ndirs) for b in range(5)}
    for d, order, a, (kx, ky, kzz) in deltas:
        idx = (kzz % nz, ky % ny, kx)
        if order == 0:
            bands[(d, 0)][idx] += a
        else:
            bands[(d, 2 * order - 1)][idx] += a
            bands[(d, 2 * order)][idx] += -1j * a
    for (d, b), v in bands.items():
        o.set_band(d, b, v)
    for d in range(ndirs):
        set_fit(o, d, k0[d], [1, 1, 1])
    o.assemble()
    out = o.result(double=(prec == 64)).astype(np.float64)
    XD, YD, ZD = 2 * nx, 2 * ny, zz * nz
    assert out.shape == (ZD, YD, XD)
    Z, Y, X = np.meshgrid(np.arange(ZD), np.arange(YD), np.arange(XD), indexing="ij")
    exp = np.zeros(out.shape)
    dxyo = np.float64(np.float32(DXY) / np.float32(2.0))
    for d, order, a, (kx, ky, kzz) in deltas:
        a = complex(np.complex64(a))
        th = 2 * np.pi * (kx * X / XD + ky * Y / YD + kzz * Z / ZD) + np.angle(a)
        if order == 0:
            exp += 2 * abs(a) * np.cos(th)
        else:
            kx0, ky0 = float(k0[d][0]) * order, float(k0[d][1]) * order
            mod = 2 * np.pi * (kx0 * (X - XD // 2) + ky0 * (Y - YD // 2)) * dxyo
            exp += 4 * abs(a) * np.cos(th + mod)
    err = np.linalg.norm(out - exp) / np.linalg.norm(exp)
    assert err < _tol(prec), err


# ----------------------------------------------------------------- Wiener filter
def covering_weight(p, k0s, amps, kabs, z0=0, band_factor=None, edges=()):
    """sum over every band (direction d, order o in -(n-1)..n-1) whose support
    |kabs - o k0_d| < rcut covers kabs, of |modamp_{d,|o|}|^2 (order 0: 1)
    times the band's own weight factor band_factor(o, kabs - o k0_d, z0)
    (suppression or damping; 1 when None), and the smallest distance of kabs
    from any band's support edge or from a circle of radius r in `edges`
    around any band's centre.  kabs: [..., 2] in 1/um."""
    rc = rcut_of(p)
    n = p.nphases // 2 + 1
    tot = np.zeros(kabs.shape[:-1])
    margin = np.full(kabs.shape[:-1], np.inf)
    for d, k0 in enumerate(k0s):
        for o2 in range(-(n - 1), n):
            dv = np.stack([kabs[..., 0] - o2 * k0[0], kabs[..., 1] - o2 * k0[1]], -1)
            dist = np.hypot(dv[..., 0], dv[..., 1])
            w = 1.0 if o2 == 0 else abs(amps[d][abs(o2)]) ** 2
            if band_factor is not None:
                w = w * band_factor(o2, dv, z0)
            tot += np.where(dist < rc, w, 0.0)
            for r in (rc,) + tuple(edges):
                margin = np.minimum(margin, np.abs(dist - r))
    return tot, margin


def grid_k(p):
    """(kx, ky) in 1/um of every stored half-spectrum element [iy][ix] (rows
    above ny/2 are negative frequencies)."""
    nxh = p.nx // 2 + 1
    iy = np.arange(p.ny)
    ys = np.where(iy < p.ny // 2, iy, iy - p.ny)
    KX = np.arange(nxh)[None, :] / (p.nx * DXY) * np.ones((p.ny, 1))
    KY = ys[:, None] / (p.ny * DXY) * np.ones((1, nxh))
    return np.stack([KX, KY], -1)


def check_filter(o, p, k0s, amps, d, bands_in, band_factor=None, apod=None, margin=0.01, zd=None, edges=()):
    """Compares every band of direction d after filterbands with the closed
    form of a generalised Wiener filter with a unit OTF: order 0 becomes
    v f(k) A(k) / D(k); a side band's positive half (re + i im) becomes
    v conj(a) f(k) A(k + o k0) / D(k + o k0) and its negative half
    (re - i im) v a f(-k) A(-k + o k0) / D(-k + o k0), where D is the
    covering weight plus wiener^2, f the band's own weight factor
    (band_factor) and A the output apodization (apod; both 1 when None).
    Both halves are zero outside the band's own support |k| <= rcut, and
    planes beyond the axial support zd[order] are zero.  Elements within
    `margin` of a support edge or of a radius in `edges` (where a factor
    switches) are left out: the float comparison may fall either way there."""
    n = p.nphases // 2 + 1
    w2 = float(np.float32(p.wiener)) ** 2
    rc = rcut_of(p)
    K = grid_k(p)
    kmag = np.hypot(K[..., 0], K[..., 1])
    one = np.ones(K.shape[:-1])
    bf = (lambda o_, dv, z: one) if band_factor is None else band_factor
    ap = (lambda o_, k, kabs, z: one) if apod is None else apod
    checked = 0
    for order in range(n):
        k0 = np.array(k0s[d], np.float64) * order
        for iz in range(p.nz):
            z0 = iz if iz <= p.nz // 2 else iz - p.nz
            keep = (kmag <= rc) & (zd is None or abs(z0) <= zd[order])
            if order == 0:
                got = o.band(d, 0, double=True)[iz]
                src = bands_in[(d, 0)][iz].astype(np.complex128)
                cov, mg = covering_weight(p, k0s, amps, K, z0, band_factor, edges)
                exp = np.where(keep, src * bf(0, K, z0) * ap(0, K, K, z0) / (cov + w2), 0.0)
                ok = (np.abs(kmag - rc) > margin) & (mg > margin)
            else:
                re = o.band(d, 2 * order - 1, double=True)[iz]
                im = o.band(d, 2 * order, double=True)[iz]
                re0 = bands_in[(d, 2 * order - 1)][iz].astype(np.complex128)
                im0 = bands_in[(d, 2 * order)][iz].astype(np.complex128)
                a = complex(np.complex64(amps[d][order]))
                kp, km = K + k0, -K + k0
                covp, mgp = covering_weight(p, k0s, amps, kp, z0, band_factor, edges)
                covm, mgm = covering_weight(p, k0s, amps, km, -z0, band_factor, edges)
                plus = np.where(keep, (re0 + 1j * im0) * np.conj(a) * bf(order, K, z0) * ap(order, K, kp, z0) /
                                (covp + w2), 0.0)
                minus = np.where(keep, (re0 - 1j * im0) * a * bf(order, -K, -z0) * ap(order, -K, km, -z0) /
                                 (covm + w2), 0.0)
                exp = np.stack([plus, minus])
                got = np.stack([re + 1j * im, re - 1j * im])
                ok = (np.abs(kmag - rc) > margin) & (mgp > margin) & (mgm > margin)
                ok[:, 0] = False  # x = 0 column: visited once, its negative half is the mirror's
                ok = np.stack([ok, ok])
            scale = max(np.abs(exp[ok]).max(), 1.0)  # inputs are unit normal; a plane may be all but zero
            err = np.abs(got[ok] - exp[ok]).max() / scale
            assert err < (2e-5 if o.prec == 32 else 1e-10), (d, order, iz, err)
            checked += int(ok.sum())
    return checked


def random_bands(p, seed):
    rng = np.random.default_rng(seed)
    nxh = p.nx // 2 + 1
    shape = (p.nz, p.ny, nxh)
    return {(d, b): (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
            for d in range(p.ndirs) for b in range(p.nphases)}


def run_filter(p, k0s, amps, prec, seed=1):
    o = ob.Oracle(p, unit_otf(p), prec)
    bands = random_bands(p, seed)
    for (d, b), v in bands.items():
        o.set_band(d, b, v)
    for d in range(p.ndirs):
        set_fit(o, d, k0s[d], amps[d])
    for d in range(p.ndirs):
        o.filter_dir(d)
    return o, bands


K0S = [(0.97, 0.31), (-0.42, 0.95), (-0.61, -0.83)]  # |k0| about 1 1/um: side bands overlap order 0


def amps_for(n):
    base = [[1, 0.6 * np.exp(0.4j), 0.35 * np.exp(-0.9j)], [1, 0.5 * np.exp(-1.3j), 0.45 * np.exp(2.2j)],
            [1, 0.7 * np.exp(2.9j), 0.25 * np.exp(0.1j)]]
    return [[complex(np.complex64(v)) for v in row[:n]] for row in base]


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("nphases", [3, 5])
def test_wiener_denominator_counts_covering_bands(prec, nphases):
    p = pin_params(64, 64, 1, 3, nphases)
    n = nphases // 2 + 1
    k0s = [tuple(float(np.float32(v)) for v in k) for k in K0S]
    amps = amps_for(n)
    o, bands = run_filter(p, k0s, amps, prec)
    checked = sum(check_filter(o, p, k0s, amps, d, bands) for d in range(3))
    assert checked > 3000


def axial_supports(p):
    """zd of each order: ceil((1 - cos alpha) / lambda_em / dkz) for order 0,
    int(1.3 zd0) for the top order, int((1 + 1/0.88) zd0) between, clamped
    below nz/2 (the detection OTF's axial extent and its order scaling)."""
    lam = p.wave / p.nimm / 1000.0
    alpha = math.asin(p.na / p.nimm)
    dkz = 1.0 / (p.nz * p.dz)
    zd0 = math.ceil((1 - math.cos(alpha)) / lam / dkz)
    n = p.nphases // 2 + 1
    zd = [zd0] + [int((1 + 1 / 0.88) * zd0)] * (n - 2) + ([int(1.3 * zd0)] if n > 1 else [])
    return [min(z, p.nz // 2 - 1) if z >= p.nz // 2 else z for z in zd]


@pytest.mark.parametrize("prec", [32, 64])
def test_wiener_denominator_3d_axial_support(prec):
    p = pin_params(48, 48, 16, 3, 5, dz=0.5)
    zd = axial_supports(p)
    assert zd == [2, 4, 2]
    k0s = [tuple(float(np.float32(v)) for v in k) for k in K0S]
    amps = amps_for(3)
    o, bands = run_filter(p, k0s, amps, prec, seed=2)
    assert sum(check_filter(o, p, k0s, amps, d, bands, zd=zd) for d in range(3)) > 10000


def dev_suppress_closed(x):
    # 1 / (1 + 20000 / (x^6 + 20)): 1/1001 at the band centre, -> 1 far from it
    return 1.0 / (1.0 + 20000.0 / (x ** 6 + 20.0))


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("ndirs", [1, 3])
def test_suppression_at_closed_values(prec, ndirs):
    """Suppression multiplies every band's weight (order 0's too, without
    damping) within 10 pixels of the band's own centre by
    1 / (1 + 20000 / (x^6 + 20)), x the distance in pixels: 1/1001 at the
    centre, 1 beyond the radius."""
    p = pin_params(64, 64, 1, ndirs, 3, bSuppress_singularities=1)
    rc = rcut_of(p)
    dk = 1.0 / (64 * DXY)
    k0s = [tuple(float(np.float32(v)) for v in k) for k in K0S[:ndirs]]
    amps = amps_for(2)[:ndirs]
    if ndirs == 1:  # side bands clear of order 0's support near the origin
        k0s = [(float(np.float32(1.25 * rc)), 0.0)]

    def supp(o2, dv, z0):
        x = np.hypot(dv[..., 0], dv[..., 1]) / dk
        return np.where(x <= 10.0, dev_suppress_closed(x), 1.0)

    o, bands = run_filter(p, k0s, amps, prec, seed=3)
    for d in range(ndirs):
        check_filter(o, p, k0s, amps, d, bands, band_factor=supp, edges=(10 * dk,))
    if ndirs == 1:
        # the band centre: order 0 keeps (1/1001) / (1/1001 + w^2) of its value
        w2 = float(np.float32(p.wiener)) ** 2
        got = o.band(0, 0, double=True)[0, 0, 0] / complex(bands[(0, 0)][0, 0, 0])
        assert abs(got - (1 / 1001) / (1 / 1001 + w2)) < (1e-6 if prec == 32 else 1e-12)


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("nz", [16])
def test_order0_damping_at_closed_values(prec, nz):
    """dampenOrder0 weights every order-0 band by (r / rcut)^2 + |z / zd0|^3
    (r its own distance from the centre): 0 at the origin, 1/4 at half the
    cut-off radius, 1 at the axial support edge.  (3D only: with one section
    zd0 is 0 and the reference's z term is 0 / 0.)"""
    p = pin_params(64, 64, nz, 3, 3, dz=0.5, bDampenOrder0=1)
    rc = rcut_of(p)
    zd = axial_supports(p) if nz > 1 else None
    k0s = [tuple(float(np.float32(v)) for v in k) for k in K0S]
    amps = amps_for(2)

    def damp(o2, dv, z0):
        if o2 != 0:
            return np.ones(dv.shape[:-1])
        r = np.hypot(dv[..., 0], dv[..., 1])
        return (r / rc) ** 2 + (abs(z0 / zd[0]) ** 3 if nz > 1 else 0.0)

    o, bands = run_filter(p, k0s, amps, prec, seed=4)
    for d in range(3):
        check_filter(o, p, k0s, amps, d, bands, band_factor=damp, zd=zd)
    g = o.band(0, 0, double=True)
    assert g[0, 0, 0] == 0  # no weight at the origin of order 0


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("nz", [1, 16])
def test_output_apodization_at_closed_values(prec, mode, nz):
    """apodizeoutput 2 multiplies by 1 - rho, 1 by cos(pi rho / 2), with
    rho = |(|kabs| / apocutoff, z / zapocutoff)| clipped at 1, apocutoff =
    rcut + |k0 of direction 0| (norders - 1) for every direction, zapocutoff
    the order-1 axial support (2D: rho = |kabs| / apocutoff)."""
    p = pin_params(64, 64, nz, 2, 3, dz=0.5, apodizeoutput=mode)
    rc = rcut_of(p)
    k0s = [(0.97, 0.31), (-0.62, 1.15)]  # different |k0|: only direction 0's sets apocutoff
    k0s = [tuple(float(np.float32(v)) for v in k) for k in k0s]
    amps = amps_for(2)[:2]
    apo = rc + math.hypot(*k0s[0])
    zd = axial_supports(p) if nz > 1 else None
    zapo = zd[1] if nz > 1 else 0

    def apod(order, k, kabs, z0):
        rho2 = (np.hypot(kabs[..., 0], kabs[..., 1]) / apo) ** 2
        if zapo > 0:
            rho2 = rho2 + (abs(z0) / zapo) ** 2
        rho = np.minimum(np.sqrt(rho2), 1.0)
        return 1 - rho if mode == 2 else np.cos(np.pi / 2 * rho)

    o, bands = run_filter(p, k0s, amps, prec, seed=5)
    for d in range(2):
        check_filter(o, p, k0s, amps, d, bands, apod=apod, zd=zd)


# ----------------------------------------------------------------- end to end
def simulate_flat(nx, ndirs, nphases, k0_bins, mods, phase0, seed):
    """Noiseless 2D SIM images (float64) of a smooth sample under a flat OTF
    (1 inside |k| < rcut, 0 outside): image_{d,p} = A + sum_o 2 Re(m_o B_o
    e^{i o phi_p}), A = IFFT[OTF FFT[S]], B_o = IFFT[OTF FFT[S e^{2 pi i o k0.r}]],
    phi_p = 2 pi p / nphases + phase0_d.  k0 on whole frequency bins, so the
    phase ramps are periodic on the grid."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:nx, 0:nx].astype(np.float64)
    S = np.full((nx, nx), 5.0)
    for _ in range(40):
        cx, cy = rng.uniform(0, nx, 2)
        dx = (x - cx + nx / 2) % nx - nx / 2
        dy = (y - cy + nx / 2) % nx - nx / 2
        S += rng.uniform(50, 200) * np.exp(-(dx ** 2 + dy ** 2) / (2 * 1.5 ** 2))
    f = np.fft.fftfreq(nx, d=DXY)
    KX, KY = np.meshgrid(f, f)
    p0 = pin_params(nx, nx, 1, ndirs, nphases)
    otf = (np.hypot(KX, KY) < rcut_of(p0)).astype(np.float64)
    n = nphases // 2 + 1
    imgs = []
    for d in range(ndirs):
        kx, ky = k0_bins[d][0] / (nx * DXY), k0_bins[d][1] / (nx * DXY)
        A = np.fft.ifft2(otf * np.fft.fft2(S)).real
        B = [None] + [np.fft.ifft2(otf * np.fft.fft2(S * np.exp(2j * np.pi * o * (kx * x + ky * y) * DXY)))
                      for o in range(1, n)]
        for ph in range(nphases):
            phi = 2 * np.pi * ph / nphases + phase0[d]
            im = A.copy()
            for o in range(1, n):
                im += 2 * np.real(mods[d][o] * B[o] * np.exp(1j * o * phi))
            imgs.append(im)
    return S, np.array(imgs)


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("nphases", [3, 5])
def test_noiseless_flat_otf_reconstructs_sample_spectrum(prec, nphases):
    nx, ndirs = 64, 3
    n = nphases // 2 + 1
    k0_bins = [(8, 2), (-3, 7), (-5, -6)]
    mods = [[1, 0.6 * np.exp(0.4j), 0.35 * np.exp(-0.9j)], [1, 0.5 * np.exp(-1.3j), 0.45 * np.exp(2.2j)],
            [1, 0.7 * np.exp(2.9j), 0.25 * np.exp(0.1j)]]
    phase0 = [0.2, -0.7, 1.9]
    S, raw = simulate_flat(nx, ndirs, nphases, k0_bins, mods, phase0, seed=7)
    p = pin_params(nx, nx, 1, ndirs, nphases, wiener=0.05)
    o = ob.Oracle(p, unit_otf(p), prec)
    o.load(raw.astype(np.float32))
    o.preprocess_bands()
    k0s = []
    for d in range(ndirs):
        k0 = (float(np.float32(k0_bins[d][0] / (nx * DXY))), float(np.float32(k0_bins[d][1] / (nx * DXY))))
        k0s.append(k0)
        # the band of order o holds n conj(m_o) e^{-i o phase0} times the shifted sample; the
        # assembly's phase ramp is centred on the zoomed grid, which adds e^{-2 pi i o k0.(nx/2, ny/2) dxy}
        centre = np.exp(-2j * np.pi * (k0[0] * (nx // 2) + k0[1] * (nx // 2)) * DXY)
        amps = [1] + [np.conj(mods[d][o]) * np.exp(-1j * o * phase0[d]) * centre ** o for o in range(1, n)]
        set_fit(o, d, k0, amps)
    o.filter_assemble()
    out = o.result(double=(prec == 64)).astype(np.float64)[0]
    XD = 2 * nx
    got = np.fft.fft2(out)
    # expected: C W(K) S^(K) on the zoomed grid, W = cov / (cov + w^2), C = XD^2 nphases inscale
    fz = np.fft.fftfreq(XD, d=1.0 / XD)  # integer bins of the zoomed grid
    KX, KY = np.meshgrid(fz, fz)
    Kabs = np.stack([KX, KY], -1) / (nx * DXY)
    mods_abs = [[abs(v) for v in row[:n]] for row in mods]
    cov, mg = covering_weight(p, k0s, mods_abs, Kabs)
    w2 = float(np.float32(p.wiener)) ** 2
    Shat = np.fft.fft2(S)
    Spad = np.zeros((XD, XD), np.complex128)
    ix = KX.astype(int) % nx
    iy = KY.astype(int) % nx
    inband = (np.abs(KX) < nx // 2) & (np.abs(KY) < nx // 2)
    Spad[inband] = Shat[iy[inband], ix[inband]]
    assert np.all(cov[~inband] == 0)  # the extended support stays inside the input grid's band
    inscale = 1.0 / (nx * nx * 1 * 2.0 * 2.0 * 1 * ndirs)
    exp = XD * XD * nphases * inscale * Spad * cov / (cov + w2)
    err = np.linalg.norm(got - exp) / np.linalg.norm(exp)
    # FP64: the stack enters as float32, which alone moves the spectrum by ~1e-7
    assert err < (2e-5 if prec == 32 else 1e-6), err
    # nothing outside the extended support, and the side bands reach beyond order 0's
    assert np.abs(got[cov == 0]).max() < 1e-4 * np.abs(got).max()
    assert (cov > 0).sum() > 2 * (np.hypot(KX, KY) < rcut_of(p) * nx * DXY).sum()
