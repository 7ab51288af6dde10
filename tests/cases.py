"""Shared inputs for the tests: the reference's test_data fixture and seeded synthetic stacks."""
import functools
import os

import numpy as np

from cudasirecon_amd import mrc, synth
from cudasirecon_amd._abi import default_params, set_params

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel_l2(a, b):
    """||a - b|| / ||b|| in float64 / complex128 (complex arrays keep their imaginary parts)."""
    a = np.asarray(a)
    b = np.asarray(b)
    t = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a = a.astype(t)
    b = b.astype(t)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


@functools.lru_cache(maxsize=None)
def load_testdata():
    """raw.dv / otf.dv with the parameters of test_data/config (MRC mode).

    Header use follows loadHeader (src/cudaSirecon/mrc.h:18-45) and
    determine_otf_dimensions for a 3D RA OTF (src/cudaSirecon/cudaSirecon.cpp:1504-1509).
    """
    h, raw = mrc.read(os.path.join(GOLDEN, "raw.dv"))
    ho, otf = mrc.read(os.path.join(GOLDEN, "otf.dv"))
    p = default_params()
    ndirs, nphases = 3, 5
    nz = h["nz"] // (h["num_waves"] * h["num_times"]) // (ndirs * nphases)
    set_params(p, nx=h["nx"], ny=h["ny"], nz=nz, dxy=h["ylen"], dz=h["zlen"], wave=h["iwav1"], ntimes=1,
               nimm=1.515, constbkgd=0.0, wiener=0.001, k0angles=[-0.8043, -1.8555, 0.2388], linespacing=0.2035,
               ndirs=ndirs, nphases=nphases, na=1.42, bRadAvgOTF=1, bDampenOrder0=1,
               nzotf=ho["nx"], nxotf=ho["ny"], nyotf=1, dkzotf=ho["xlen"], dkrotf=ho["ylen"])
    return p, otf.astype(np.complex64), raw.astype(np.float32)


@functools.lru_cache(maxsize=None)
def synthetic(nx, ny, nz, nphases=5, seed=11, photons=5000.0, **kw):
    p, tab, raw, truth = synth.make_dataset(nx, ny, nz, nphases=nphases, seed=seed, photons=photons, **kw)
    return p, tab, raw, truth


def full_otf_from_ra(p, tab):
    """The full (not rotationally averaged) OTF table of an RA one: [order][kz][ky][kx] over
    kx >= 0, ky in FFT order, each entry the RA profile linearly interpolated at
    kr = |(kx, ky)| (clamped at the table's edge), with the same kr and kz steps.
    Returns (params with bRadAvgOTF = 0 and the full geometry, table)."""
    tab = np.asarray(tab)
    norders, nxotf, nzotf = tab.shape
    nyf = 2 * (nxotf - 1)
    kx = np.arange(nxotf, dtype=np.float64)
    ky = np.fft.fftfreq(nyf, 1.0 / nyf)
    r = np.minimum(np.hypot(kx[None, :], ky[:, None]), nxotf - 1)  # [ky][kx]
    i0 = np.minimum(np.floor(r).astype(int), nxotf - 2)
    a = r - i0
    full = np.empty((norders, nzotf, nyf, nxotf), np.complex64)
    for o in range(norders):
        t = tab[o].astype(np.complex128)  # [kr][kz]
        v = t[i0] * (1 - a)[..., None] + t[i0 + 1] * a[..., None]  # [ky][kx][kz]
        full[o] = np.transpose(v, (2, 0, 1))
    q = p.copy()
    q.bRadAvgOTF = 0
    q.nxotf, q.nyotf, q.nzotf = nxotf, nyf, nzotf
    return q, full


def tiff_frames(path):
    """All pages of a TIFF as one array (Pillow's libtiff-backed reader)."""
    from PIL import Image

    im = Image.open(path)
    out = []
    for i in range(getattr(im, "n_frames", 1)):
        im.seek(i)
        out.append(np.array(im))
    return np.stack(out)


@functools.lru_cache(maxsize=None)
def load_testdata_tiff():
    """raw.tif / otf.tif with test_data/config-tiff, as the TIFF mode derives them.

    dxy/dz/wave come from xyres/zres/wavelength (wavelength keeps its default
    530); the RA OTF page is [kr][kz] complex with nzotf = width/2, nxotf =
    height, dkzotf = 1/(zresPSF*nzotf), dkrotf = 1/(dxy*(nxotf-1)*2)
    (src/cudaSirecon/cudaSirecon.cpp:1449-1470, 1540-1571).
    """
    raw = tiff_frames(os.path.join(GOLDEN, "raw.tif")).astype(np.float32)
    otf_pages = tiff_frames(os.path.join(GOLDEN, "otf.tif")).astype(np.float32)
    otf = np.ascontiguousarray(otf_pages).view(np.complex64)  # [order][kr][kz]
    nzotf, nxotf = otf.shape[2], otf.shape[1]
    ndirs, nphases = 3, 5
    nz = raw.shape[0] // (ndirs * nphases)
    dxy, dzpsf = np.float32(0.08), np.float32(0.125)
    p = default_params()
    set_params(p, nx=raw.shape[2], ny=raw.shape[1], nz=nz, dxy=0.08, dz=0.125, wave=530, ntimes=1,
               nimm=1.515, constbkgd=0.0, wiener=0.001, k0angles=[0.8043, 1.8555, -0.2388], linespacing=0.2035,
               ndirs=ndirs, nphases=nphases, na=1.42, bRadAvgOTF=1, bDampenOrder0=1, bFastSIM=1,
               nzotf=nzotf, nxotf=nxotf, nyotf=1,
               dkzotf=float(np.float32(1) / (dzpsf * np.float32(nzotf))),
               dkrotf=float(np.float32(1) / (dxy * np.float32(nxotf - 1) * np.float32(2))))
    return p, np.ascontiguousarray(otf[:3]), raw
