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
