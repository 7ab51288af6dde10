"""simr_fft3d_dev_noalloc (the FFT entry point that allocates nothing) against
simr_fft3d_dev on the same device arrays, through the C ABI.

Two calls in a row, the second on a smaller array (both with a length that
runs as a direct DFT through the scratch array, so the second call reuses
what the first left there), forward and inverse: the two entry points agree
to the FFT tolerance (rel-L2 2e-6) and both match numpy.  An array larger
than the scratch is refused before any device work.
"""
import ctypes as C

import numpy as np
import pytest

from cases import rel_l2
from cudasirecon_amd import engine

pytestmark = pytest.mark.gpu


def _lib():
    L = engine.lib()
    for name in ("simr_fft3d_dev", "simr_fft3d_dev_noalloc"):
        f = getattr(L, name)
        f.restype = C.c_int
        f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
    L.simr_fft3d_noalloc_max_elems.restype = C.c_size_t
    L.simr_fft3d_noalloc_max_elems.argtypes = []
    return L


def _run(L, fn, x, inverse):
    import torch

    nz, ny, nx = x.shape
    d = torch.from_numpy(x.view(np.float32)).to("cuda")
    torch.cuda.synchronize()
    rc = getattr(L, fn)(d.data_ptr(), nx, ny, nz, int(inverse), None)
    assert rc == 0, (fn, rc, L.simr_last_error())
    torch.cuda.synchronize()
    return d.cpu().numpy().view(np.complex64).reshape(x.shape)


@pytest.mark.parametrize("inverse", [False, True])
def test_noalloc_fft_matches_fft3d_dev(inverse):
    L = _lib()
    rng = np.random.default_rng(5 + inverse)
    for shape in [(13, 22, 60), (7, 10, 11)]:  # (nz, ny, nx): 13, 22 and 11 have no transform kernel
        print("case: noalloc vs fft3d_dev, shape %s, inverse %d" % (shape, inverse), flush=True)
        x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
        a = _run(L, "simr_fft3d_dev", x, inverse)
        b = _run(L, "simr_fft3d_dev_noalloc", x, inverse)
        x64 = x.astype(np.complex128)
        ref = np.fft.ifftn(x64) * x.size if inverse else np.fft.fftn(x64)
        assert rel_l2(b, a) < 2e-6, (shape, rel_l2(b, a))
        assert rel_l2(a, ref) < 2e-6 and rel_l2(b, ref) < 2e-6, (shape, rel_l2(a, ref), rel_l2(b, ref))


def test_noalloc_fft_refuses_arrays_beyond_its_scratch():
    import torch

    L = _lib()
    cap = L.simr_fft3d_noalloc_max_elems()
    assert cap == ((64 << 20) - 3 * 4096 * 8) // 8
    d = torch.zeros(2, dtype=torch.float32, device="cuda")  # never touched: the size check comes first
    rc = L.simr_fft3d_dev_noalloc(d.data_ptr(), 4093, 2048, 1, 0, None)  # 4093 is prime: needs scratch
    assert rc == -1 and b"scratch" in L.simr_last_error()
    torch.cuda.synchronize()
