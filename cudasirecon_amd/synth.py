"""Deterministic synthetic 2D/3D SIM data: sample, OTF table and raw stacks.

Used by the tests and the benchmark (there is no network, so no measured
datasets beyond the two small fixtures the reference ships).

Forward model (matches the band convention the reconstruction assumes,
src/cudaSirecon/cudaSirecon.cpp:606-617 and gpuFunctionsImpl.cu:1741-1951):
    image_{d,p}(r) = A(r) + sum_{o>=1} 2 Re( m_o B_o(r) e^{i o phi_{d,p}} )
    A   = IFFT[ OTF_0 . FFT[S] ]
    B_o = IFFT[ OTF_o . FFT[S e^{+2 pi i o k0_d . r}] ]
with phi_{d,p} = 2 pi p / nphases + phase0_d, so that the separated band
bandre + i bandim of order o carries OTF_o(k) S(k - o k0) (the sign the
reconstruction's findk0 / fitk0andmodamps recover).  The OTF of every order is a
smooth rotationally symmetric model sampled on the reconstruction's RA table
grid ([kr][kz], kz wrapped), so data and reconstruction share one OTF.
"""
import math

import numpy as np

try:
    import scipy.fft as _fft

    def _fftn(a):
        return _fft.fftn(a, workers=-1)

    def _ifftn(a):
        return _fft.ifftn(a, workers=-1)
except Exception:  # pragma: no cover
    _fftn, _ifftn = np.fft.fftn, np.fft.ifftn

from ._abi import default_params, set_params


def otf_model(order, kr, kz, na, nimm, wave_nm):
    """Rotationally symmetric OTF model of one order at (kr, kz) in 1/um."""
    lam = wave_nm / 1000.0
    kc = 2 * na / lam
    lem = lam / nimm
    alpha = math.asin(na / nimm)
    kzmax = (1 - math.cos(alpha)) / lem
    w = np.clip(1.0 - kr / kc, 0.0, None)
    lat = w ** 1.5
    if order == 0:
        ax = np.exp(-(kz / (0.9 * kzmax)) ** 2)
    elif order == 1:
        ax = 0.5 * (np.exp(-((kz - 0.6 * kzmax) / (0.7 * kzmax)) ** 2) +
                    np.exp(-((kz + 0.6 * kzmax) / (0.7 * kzmax)) ** 2))
    else:
        ax = np.exp(-(kz / (0.8 * kzmax)) ** 2)
    return (lat * ax).astype(np.float64)


def make_otf_table(norders, na, nimm, wave_nm, dxy, nzotf=64, dzpsf=0.125, nxotf=None, two_d=False):
    """RA OTF table [order][kr][kz] (complex64) plus its geometry (nxotf, nzotf, dkr, dkz)."""
    if nxotf is None:
        nxotf = 129
    dkr = 1.0 / (dxy * (nxotf - 1) * 2)
    if two_d:
        nzotf = 1
        dkz = 1.0
    else:
        dkz = 1.0 / (dzpsf * nzotf)
    kr = np.arange(nxotf) * dkr
    iz = np.arange(nzotf)
    kzi = np.where(iz < (nzotf + 1) // 2, iz, iz - nzotf) * dkz
    tab = np.zeros((norders, nxotf, nzotf), np.complex64)
    for o in range(norders):
        v = otf_model(o, kr[:, None], 0.0 * kzi[None, :] if two_d else kzi[None, :], na, nimm, wave_nm)
        tab[o] = v
    tab /= tab[0, 0, 0]
    return tab, dict(nxotf=nxotf, nzotf=nzotf, nyotf=1, dkrotf=dkr, dkzotf=dkz)


def make_sample(nx, ny, nz, seed, nbeads=400, nfil=20, sigma=1.0, amp=(200.0, 2000.0)):
    """Gaussian beads plus random-walk filaments, float32 [nz][ny][nx]."""
    rng = np.random.default_rng(seed)
    s = np.zeros((nz, ny, nx), np.float32)
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r = int(math.ceil(4 * sigma))
    for _ in range(nbeads):
        cz = rng.uniform(0, nz) if nz > 1 else 0.0
        cy, cx = rng.uniform(0, ny), rng.uniform(0, nx)
        a = rng.uniform(*amp)
        iz0 = int(cz)
        iy0, ix0 = int(cy), int(cx)
        zs = slice(max(iz0 - r, 0), min(iz0 + r + 1, nz))
        ys = slice(max(iy0 - r, 0), min(iy0 + r + 1, ny))
        xs = slice(max(ix0 - r, 0), min(ix0 + r + 1, nx))
        d2 = (xx[zs, ys, xs] - cx) ** 2 + (yy[zs, ys, xs] - cy) ** 2
        if nz > 1:
            d2 = d2 + (zz[zs, ys, xs] - cz) ** 2
        s[zs, ys, xs] += a * np.exp(-d2 / (2 * sigma ** 2))
    for _ in range(nfil):
        p = np.array([rng.uniform(0, nz) if nz > 1 else 0.0, rng.uniform(0, ny), rng.uniform(0, nx)])
        d = rng.normal(size=3)
        if nz == 1:
            d[0] = 0
        d /= np.linalg.norm(d) + 1e-12
        a = rng.uniform(*amp) * 0.3
        for _ in range(int(1.5 * max(nx, ny))):
            d += 0.15 * rng.normal(size=3)
            if nz == 1:
                d[0] = 0
            else:
                d[0] *= 0.3
            d /= np.linalg.norm(d) + 1e-12
            p = p + 0.7 * d
            iz, iy, ix = int(p[0]) % nz, int(p[1]) % ny, int(p[2]) % nx
            s[iz, iy, ix] += a
    return s


def simulate(sample, ndirs, nphases, k0_list, dxy, dz, otf_fn, modamps, phase0=None, seed=0,
             photons=None, background=90.0, u16=True, fastSI=False):
    """Raw SIM sections [nImages][ny][nx] in the reference's section order.

    k0_list: per-direction order-1 pattern vectors (kx, ky) in 1/um.
    otf_fn(order, kr, kz) -> OTF values; modamps: [m_1, m_2, ...].
    photons: peak expected photon count for Poisson noise (None = noiseless).
    """
    nz, ny, nx = sample.shape
    norders = nphases // 2 + 1
    rng = np.random.default_rng(seed)
    ky = np.fft.fftfreq(ny, d=dxy)
    kx = np.fft.fftfreq(nx, d=dxy)
    kz = np.fft.fftfreq(nz, d=dz) if nz > 1 else np.zeros(1)
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    KR = np.sqrt(KX ** 2 + KY ** 2)
    y = np.arange(ny)[None, :, None] * dxy
    x = np.arange(nx)[None, None, :] * dxy
    S = sample.astype(np.complex64)
    if phase0 is None:
        phase0 = [0.0] * ndirs
    imgs = np.zeros((ndirs, nz, nphases, ny, nx), np.float32)
    for d in range(ndirs):
        comps = []
        for o in range(norders):
            otf = otf_fn(o, KR, KZ).astype(np.float32)
            if o == 0:
                field = S
            else:
                ph = np.exp(2j * np.pi * o * (k0_list[d][0] * x + k0_list[d][1] * y)).astype(np.complex64)
                field = S * ph
            comps.append(_ifftn(otf * _fftn(field)).astype(np.complex64))
        for p in range(nphases):
            phi = 2 * np.pi * p / nphases + phase0[d]
            im = comps[0].real.copy()
            for o in range(1, norders):
                im += 2 * np.real(modamps[o - 1] * comps[o] * np.exp(1j * o * phi))
            imgs[d, :, p] = im
    imgs = np.clip(imgs, 0, None)
    if photons is not None:
        scale = photons / max(float(imgs.max()), 1e-12)
        imgs = rng.poisson(imgs * scale + background).astype(np.float32)
    else:
        imgs = imgs + background
    if fastSI:
        out = imgs.transpose(1, 0, 2, 3, 4).reshape(-1, ny, nx)
    else:
        out = imgs.reshape(-1, ny, nx)
    if u16:
        out = np.clip(np.rint(out), 0, 65535).astype(np.uint16)
    return out


def simulate_gpu(sample, ndirs, nphases, k0_list, dxy, dz, otf_fn, modamps, phase0=None, seed=0,
                 photons=None, background=90.0, u16=True, fastSI=False):
    """simulate() with the transforms, products and Poisson draws on the GPU (torch).

    Same forward model; the noise comes from torch's generator, so a stack
    differs from simulate()'s for the same seed.  For generating large
    benchmark stacks quickly; the tests use simulate().
    """
    import torch

    dev = "cuda"
    nz, ny, nx = sample.shape
    norders = nphases // 2 + 1
    ky = np.fft.fftfreq(ny, d=dxy)
    kx = np.fft.fftfreq(nx, d=dxy)
    kz = np.fft.fftfreq(nz, d=dz) if nz > 1 else np.zeros(1)
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    KR = np.sqrt(KX ** 2 + KY ** 2)
    otfs = [torch.from_numpy(np.ascontiguousarray(otf_fn(o, KR, KZ), dtype=np.float32)).to(dev)
            for o in range(norders)]
    del KZ, KY, KX, KR
    S = torch.from_numpy(np.ascontiguousarray(sample, dtype=np.float32)).to(dev).to(torch.complex64)
    y = torch.arange(ny, device=dev, dtype=torch.float64)[None, :, None] * dxy
    x = torch.arange(nx, device=dev, dtype=torch.float64)[None, None, :] * dxy
    if phase0 is None:
        phase0 = [0.0] * ndirs
    imgs = torch.zeros((ndirs, nz, nphases, ny, nx), dtype=torch.float32, device=dev)
    for d in range(ndirs):
        comps = []
        for o in range(norders):
            if o == 0:
                field = S
            else:
                ph = torch.exp(2j * math.pi * o * (k0_list[d][0] * x + k0_list[d][1] * y)).to(torch.complex64)
                field = S * ph
            comps.append(torch.fft.ifftn(otfs[o] * torch.fft.fftn(field)))
        for p in range(nphases):
            phi = 2 * math.pi * p / nphases + phase0[d]
            im = comps[0].real.clone()
            for o in range(1, norders):
                im += 2 * torch.real(modamps[o - 1] * comps[o] * complex(math.cos(o * phi), math.sin(o * phi)))
            imgs[d, :, p] = im
        del comps
    imgs.clamp_(min=0)
    if photons is not None:
        scale = photons / max(float(imgs.max()), 1e-12)
        g = torch.Generator(device=dev)
        g.manual_seed(int(seed))
        imgs = torch.poisson(imgs * scale + background, generator=g)
    else:
        imgs += background
    out = imgs.transpose(0, 1) if fastSI else imgs
    out = out.reshape(-1, ny, nx)
    if u16:
        out = torch.clamp(torch.round(out), 0, 65535).to(torch.int32).cpu().numpy().astype(np.uint16)
    else:
        out = out.cpu().numpy()
    torch.cuda.synchronize()
    return out


def make_dataset(nx, ny, nz, ndirs=3, nphases=5, seed=20261202, dxy=0.08, dz=0.125, wave=515, na=1.42,
                 nimm=1.515, ls=0.2035, angle0=1.648, modamps=(0.4, 0.2), photons=5000.0, background=90.0,
                 nbeads=None, nfil=20, nzotf=64, wiener=0.001, u16=True, angles=None, k0_jitter_px=0.0,
                 gpu=False, **param_overrides):
    """A complete synthetic case: (params, otf_table, raw, truth dict).

    gpu=True simulates with simulate_gpu (torch on the GPU; different noise draws)."""
    norders = nphases // 2 + 1
    two_d = nz == 1
    if angles is None:
        angles = [angle0 + i * 2 * math.pi / ndirs for i in range(ndirs)] if two_d else \
                 [angle0 + i * math.pi / ndirs for i in range(ndirs)]
    k0mag = 1.0 / ls if two_d else 1.0 / ls / (norders - 1)
    rng = np.random.default_rng(seed + 7)
    k0 = []
    for a in angles:
        jit = rng.normal(size=2) * k0_jitter_px / (nx * dxy)
        k0.append((k0mag * math.cos(a) + jit[0], k0mag * math.sin(a) + jit[1]))
    if nbeads is None:
        nbeads = max(50, int(400 * nx * ny * max(nz, 1) / (256 * 256 * 16)))
    sample = make_sample(nx, ny, nz, seed, nbeads=nbeads, nfil=nfil)
    tab, og = make_otf_table(norders, na, nimm, wave, dxy, nzotf=nzotf, two_d=two_d)
    fn = lambda o, kr, kz: otf_model(o, kr, kz, na, nimm, wave)  # noqa: E731
    ma = list(modamps)[: norders - 1]
    raw = (simulate_gpu if gpu else simulate)(sample, ndirs, nphases, k0, dxy, dz, fn, ma, seed=seed + 1,
                                              photons=photons, background=background, u16=u16)
    p = default_params()
    set_params(p, nx=nx, ny=ny, nz=nz, ndirs=ndirs, nphases=nphases, dxy=dxy, dz=dz if nz > 1 else 0.0,
               wave=wave, na=na, nimm=nimm, linespacing=ls, k0angles=list(angles), wiener=wiener,
               bRadAvgOTF=1, constbkgd=background, **og)
    if nz > 1:
        p.bDampenOrder0 = 1
    for k, v in param_overrides.items():
        set_params(p, **{k: v})
    truth = dict(k0=k0, modamps=ma, sample=sample)
    return p, tab, raw, truth


def make_psf_stack(nx, ny, nz, nphases=5, seed=0, sigma_xy=1.5, sigma_z=2.0, center=None, kz_cycles=0.08,
                   modamps=(0.8, 0.4), phase0=0.3, background=100.0, peak=10000.0, noise=True):
    """Bead PSF stack for makeotf: float32 sections [nz*nphases][ny][nx] in Z->Phase order.

    A point bead at `center` (pixels, sub-pixel offsets allowed) under a
    3-beam pattern: the detection PSF is an anisotropic Gaussian; phase j of
    order o multiplies it by cos(o*(phase0 + 2*pi*j/nphases)), and order 1
    also carries the axial modulation cos(2*pi*kz_cycles*(z - zc)).
    """
    rng = np.random.default_rng(seed)
    if center is None:
        center = (nx / 2 + 0.3, ny / 2 - 0.4, nz / 2 + 0.2)
    xc, yc, zc = center
    x = np.arange(nx)[None, None, :]
    y = np.arange(ny)[None, :, None]
    z = np.arange(nz)[:, None, None]
    h = np.exp(-((x - xc) ** 2 + (y - yc) ** 2) / (2 * sigma_xy ** 2) - (z - zc) ** 2 / (2 * sigma_z ** 2))
    axial = np.cos(2 * np.pi * kz_cycles * (z - zc))
    norders = nphases // 2 + 1
    out = np.empty((nz, nphases, ny, nx), np.float64)
    for j in range(nphases):
        ph = phase0 + 2 * np.pi * j / nphases
        pat = np.ones((nz, 1, 1))
        if norders > 1:
            pat = pat + modamps[0] * axial * np.cos(ph)
        if norders > 2:
            pat = pat + modamps[1] * np.cos(2 * ph)
        out[:, j] = background + peak * h * pat
    if noise:
        out = rng.poisson(np.clip(out, 0, None)).astype(np.float64)
    return out.reshape(nz * nphases, ny, nx).astype(np.float32)
