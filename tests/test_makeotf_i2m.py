"""makeotf -I2M: the oracle's I2M OTF (src/otf/radialft.cpp:267-286, 624-664,
306-307, 345-357, 383-386, 425-430) pinned on the CPU.

Pins:
  * a point bead on a flat background: its OTF is 1 at every (kr, kz) after
    the order-0 rescale, and the centre is the bead's pixel;
  * the I2M branch equals the main branch run with one phase, no
    apodization, no bead-size compensation, no modify, and the section nz/2
    border estimate as every section's background;
  * the camera fix acts only on an exact -1 in the last pixel.
"""
import numpy as np
import pytest

from cudasirecon_amd._abi import otf_default_params, set_otf_params
from cudasirecon_amd import synth
from oracle import binding as ob


def params(**kw):
    p = otf_default_params()
    set_otf_params(p, dr=0.08, dz=0.125, linespacing=0.2035, k0angle=-0.8, na=1.42, nimm=1.515, wavelength=515)
    return set_otf_params(p, **kw)


def border_mean(sec, border=20):
    ny, nx = sec.shape
    y = np.arange(ny)[:, None]
    x = np.arange(nx)[None, :]
    m = (y < border) | (y > ny - border) | (x < border) | (x > nx - border)
    return np.float32(np.sum(sec[m].astype(np.float64)) / m.sum())


@pytest.mark.parametrize("shape,pos", [((64, 64, 16), (30, 33, 7)), ((48, 48, 9), (25, 20, 4)), ((64, 64, 1), (31, 31, 0))])
def test_point_bead_gives_flat_otf(shape, pos):
    nx, ny, nz = shape
    vol = np.full((nz, ny, nx), 100.0, np.float32)
    x, y, z = pos
    vol[z, y, x] = 1100.0
    otf, info = ob.makeotf_i2m(params(), vol)
    assert otf.shape == (min(nx, ny) // 2 + 1, nz)
    assert (info.xc, info.yc, info.zc) == (x, y, z)
    np.testing.assert_allclose(otf.real, 1.0, atol=2e-5)
    np.testing.assert_allclose(otf.imag, 0.0, atol=2e-5)


@pytest.mark.parametrize("kx", [0, 3])
def test_i2m_is_the_single_phase_main_path(kx):
    nx, ny, nz = 64, 64, 12
    vol = synth.make_psf_stack(nx, ny, nz, 1, seed=5, center=(30.4, 33.7, 5.6))
    p = params(interpkr=[kx, 12 if kx else 0])
    got, ginfo = ob.makeotf_i2m(p, vol)
    bg = np.full(nz, border_mean(vol[nz // 2]), np.float32)
    q = params(nphases=1, napodize=0, do_compen=0, ifixz=0, ifixr=0, interpkr=[kx, 12 if kx else 0])
    ref, rinfo, _ = ob.makeotf(q, vol, section_bg=bg)
    np.testing.assert_array_equal(got, ref[0])
    assert (ginfo.xc, ginfo.yc, ginfo.zc) == (rinfo.xc, rinfo.yc, rinfo.zc)


def test_camera_fix_only_for_minus_one():
    nx, ny, nz = 48, 48, 6
    vol = synth.make_psf_stack(nx, ny, nz, 1, seed=3)
    a = vol.copy()
    a[:, -1, -1] = -1.0
    b = vol.copy()
    b[:, -1, -1] = vol[:, -1, -2]
    np.testing.assert_array_equal(ob.makeotf_i2m(params(), a)[0], ob.makeotf_i2m(params(), b)[0])
    c = vol.copy()
    c[:, -1, -1] = -5.0  # not the camera's code: kept (and so in the border background)
    assert not np.array_equal(ob.makeotf_i2m(params(), c)[0], ob.makeotf_i2m(params(), b)[0])
