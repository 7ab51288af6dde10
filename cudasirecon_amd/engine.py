"""ctypes binding of libcudasirecon.so's stage-level ABI (include/simr.h).

The shared library is the product; this module is a thin host-side mirror of
it for Python callers and tests.  It raises if the library is missing: there
is no fallback path.
"""
import ctypes as C
import os

import numpy as np

from ._abi import (SimrFit, SimrOtfInfo, SimrOtfParams, SimrParams, SimrRequest, SimrRequestInfo, SimrServerStats,
                   norders_of, output_shape)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcudasirecon.so")
_lib = None


class SimrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("simr error %d: %s" % (code, msg))
        self.code = code


def lib():
    """Load libcudasirecon.so (built by __graft_entry__.build() / `make -C cudasirecon_amd/csrc`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libcudasirecon.so not built: run `make -C cudasirecon_amd/csrc` (%s)" % LIB_PATH)
    L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
    vp, ip, fp, dp = C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_double)
    P = C.POINTER(SimrParams)
    F = C.POINTER(SimrFit)
    sigs = {
        "simr_params_default": (None, [P]),
        "simr_last_error": (C.c_char_p, []),
        "simr_set_last_error": (None, [C.c_char_p]),
        "simr_version": (C.c_char_p, []),
        "simr_params_size": (C.c_size_t, []),
        "simr_fit_size": (C.c_size_t, []),
        "simr_otf_size": (C.c_size_t, [P]),
        "simr_otf_total": (C.c_size_t, [P]),
        "simr_footprint": (C.c_size_t, [P]),
        "simr_create": (ip, [P, fp, ip, C.POINTER(vp)]),
        "simr_destroy": (ip, [vp]),
        "simr_get_params": (ip, [vp, P]),
        "simr_set_raw_f32_host": (ip, [vp, fp]),
        "simr_set_raw_u16_host": (ip, [vp, C.POINTER(C.c_uint16)]),
        "simr_set_raw_f32_dev": (ip, [vp, vp]),
        "simr_set_raw_u16_dev": (ip, [vp, vp]),
        "simr_set_flatfield": (ip, [vp, fp, fp]),
        "simr_load_and_rescale": (ip, [vp, ip]),
        "simr_set_time_index": (ip, [vp, ip]),
        "simr_find_modulation": (ip, [vp]),
        "simr_filter_assemble": (ip, [vp]),
        "simr_process_volume": (ip, [vp]),
        "simr_result_elems": (C.c_size_t, [vp]),
        "simr_get_result_host": (ip, [vp, fp]),
        "simr_result_dev": (vp, [vp]),
        "simr_get_fit": (ip, [vp, ip, F]),
        "simr_set_fit": (ip, [vp, ip, F]),
        "simr_get_band_host": (ip, [vp, ip, ip, fp]),
        "simr_set_band_host": (ip, [vp, ip, ip, fp]),
        "simr_filter_dir": (ip, [vp, ip]),
        "simr_debug_overlaps": (ip, [vp, ip, ip, ip, C.c_float, C.c_float, dp, dp, dp]),
        "simr_debug_overlap_volumes": (ip, [vp, ip, ip, ip, C.c_float, C.c_float, fp, fp]),
        "simr_sync": (ip, [vp]),
        "simr_stream": (vp, [vp]),
        "simr_enable_timing": (ip, [vp, ip]),
        "simr_time_stage": (ip, [vp, C.c_char_p]),
        "simr_stage_timings": (ip, [vp, C.POINTER(C.c_char_p), fp, ip]),
        "simr_fft3d_dev": (ip, [vp, ip, ip, ip, ip, vp]),
        "simr_fft3d_host": (ip, [fp, ip, ip, ip, ip]),
        "simr_host_alloc": (C.c_void_p, [C.c_size_t]),
        "simr_host_free": (ip, [vp]),
        "simr_fft_size_supported": (ip, [ip]),
        "simr_update_params": (ip, [vp, P]),
        "simr_otf_params_default": (None, [C.POINTER(SimrOtfParams)]),
        "simr_otf_elems": (C.c_size_t, [C.POINTER(SimrOtfParams), ip, ip, ip]),
        "simr_makeotf": (ip, [C.POINTER(SimrOtfParams), fp, ip, ip, ip, fp, fp, fp, C.POINTER(SimrOtfInfo)]),
        "simr_makeotf_i2m": (ip, [C.POINTER(SimrOtfParams), fp, ip, ip, ip, fp, C.POINTER(SimrOtfInfo)]),
        "simr_get_result_dev": (ip, [vp, vp]),
        "simr_time0_state_size": (C.c_size_t, [vp]),
        "simr_get_time0_state": (ip, [vp, vp, C.c_size_t]),
        "simr_set_time0_state": (ip, [vp, vp, C.c_size_t]),
        "simr_server_create": (ip, [ip, C.c_size_t, ip, C.POINTER(vp)]),
        "simr_server_add_setup": (ip, [vp, P, fp, C.POINTER(ip)]),
        "simr_server_submit": (ip, [vp, C.POINTER(SimrRequest), C.POINTER(C.c_int64)]),
        "simr_server_wait": (ip, [vp, C.c_int64, C.POINTER(SimrRequestInfo)]),
        "simr_server_stats_get": (ip, [vp, C.POINTER(SimrServerStats)]),
        "simr_server_destroy": (ip, [vp]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc < 0:
        raise SimrError(rc, lib().simr_last_error().decode())
    return rc


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def default_params():
    p = SimrParams()
    lib().simr_params_default(C.byref(p))
    return p


def otf_default_params():
    p = SimrOtfParams()
    lib().simr_otf_params_default(C.byref(p))
    return p


def makeotf(params, psf, section_bg=None, want_band0=False):
    """PSF sections [nz*nphases][ny][nx] (Z->Phase order) -> (otf [nbands][nkr][nz] complex64, info, band0)."""
    psf = np.ascontiguousarray(psf, dtype=np.float32)
    nsec, ny, nx = psf.shape
    n = lib().simr_otf_elems(C.byref(params), nx, ny, nsec)
    if n == 0:
        raise SimrError(-1, "invalid makeotf geometry")
    out = np.empty(2 * n, np.float32)
    bg = None if section_bg is None else np.ascontiguousarray(section_bg, dtype=np.float32)
    nz = nsec // params.nphases
    band0 = np.empty((nz, ny, nx), np.float32) if want_band0 else None
    info = SimrOtfInfo()
    _check(lib().simr_makeotf(C.byref(params), _fp(psf), nx, ny, nsec, None if bg is None else _fp(bg), _fp(out),
                              None if band0 is None else _fp(band0), C.byref(info)))
    otf = out.view(np.complex64).reshape(info.nbands_out, info.nkr, info.nz)
    return otf, info, band0


def makeotf_i2m(params, i2m):
    """I2M sections [nz][ny][nx] -> (I2M OTF [nkr][nz] complex64, info) (makeotf -I2M)."""
    i2m = np.ascontiguousarray(i2m, dtype=np.float32)
    nz, ny, nx = i2m.shape
    nkr = min(nx, ny) // 2 + 1
    out = np.empty(2 * nkr * nz, np.float32)
    info = SimrOtfInfo()
    _check(lib().simr_makeotf_i2m(C.byref(params), _fp(i2m), nx, ny, nz, _fp(out), C.byref(info)))
    return out.view(np.complex64).reshape(nkr, nz), info


class _PinnedBlock:
    """Page-locked host memory from simr_host_alloc, freed when the last array on it goes."""

    def __init__(self, nbytes):
        ptr = lib().simr_host_alloc(nbytes)
        if not ptr:
            raise SimrError(-1, "simr_host_alloc(%d): %s" % (nbytes, lib().simr_last_error().decode()))
        self.ptr = ptr

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().simr_host_free(C.c_void_p(self.ptr))
            self.ptr = None


def pinned_empty(shape, dtype=np.float32):
    """A numpy array in page-locked host memory: stacks and results in such arrays move
    between host and GPU by DMA at the link's rate (simr_host_alloc)."""
    dt = np.dtype(dtype)
    n = max(int(np.prod(shape)) * dt.itemsize, 1)
    block = _PinnedBlock(n)
    buf = (C.c_char * n).from_address(block.ptr)
    buf._block = block  # the buffer keeps the block, the array keeps the buffer
    return np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape)


def footprint(params):
    return lib().simr_footprint(C.byref(params))


def fft3d(a, inverse=False):
    """The engine's FFT on the current GPU: complex64 array with 1-3 dims (last axis fastest)."""
    b = np.ascontiguousarray(a, dtype=np.complex64).copy()
    dims = list(b.shape[::-1]) + [1] * (3 - b.ndim)
    _check(lib().simr_fft3d_host(_fp(b.view(np.float32)), dims[0], dims[1], dims[2], int(bool(inverse))))
    return b


class Reconstructor:
    """One reconstruction context on one GPU (owns a HIP stream and a device arena)."""

    def __init__(self, params, otf, device=0):
        self.p = params.copy()
        self.norders = norders_of(self.p)
        otf = np.ascontiguousarray(otf, dtype=np.complex64)
        n_dirs_otf = self.p.ndirs if self.p.bOneOTFperAngle else 1
        need = n_dirs_otf * self.norders * lib().simr_otf_size(C.byref(self.p))
        if otf.size < need:
            raise ValueError("OTF has %d complex entries, need %d" % (otf.size, need))
        self._otf = otf
        h = C.c_void_p()
        _check(lib().simr_create(C.byref(self.p), _fp(otf.view(np.float32)), device, C.byref(h)))
        self.h = h
        self.H = (self.p.nx // 2 + 1) * self.p.ny * self.p.nz

    def close(self):
        if getattr(self, "h", None):
            lib().simr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- input
    def set_raw(self, raw):
        raw = np.ascontiguousarray(raw)
        n = self.p.nz * self.p.ndirs * self.p.nphases * self.p.ny * self.p.nx
        if raw.size != n:
            raise ValueError("raw has %d pixels, expected %d" % (raw.size, n))
        if raw.dtype == np.uint16:
            _check(lib().simr_set_raw_u16_host(self.h, raw.ctypes.data_as(C.POINTER(C.c_uint16))))
        else:
            raw = np.ascontiguousarray(raw, dtype=np.float32)
            _check(lib().simr_set_raw_f32_host(self.h, _fp(raw)))

    def set_raw_device(self, ptr, u16=True):
        if u16:
            _check(lib().simr_set_raw_u16_dev(self.h, C.c_void_p(ptr)))
        else:
            _check(lib().simr_set_raw_f32_dev(self.h, C.c_void_p(ptr)))

    def set_flatfield(self, background=None, slope=None):
        bg = None if background is None else np.ascontiguousarray(background, dtype=np.float32)
        sl = None if slope is None else np.ascontiguousarray(slope, dtype=np.float32)
        _check(lib().simr_set_flatfield(self.h, None if bg is None else _fp(bg), None if sl is None else _fp(sl)))

    # -- stages
    def load_and_rescale(self, it=0):
        _check(lib().simr_load_and_rescale(self.h, it))

    def set_time(self, it):
        _check(lib().simr_set_time_index(self.h, it))

    def find_modulation(self):
        _check(lib().simr_find_modulation(self.h))

    def filter_assemble(self):
        _check(lib().simr_filter_assemble(self.h))

    def process_volume(self):
        _check(lib().simr_process_volume(self.h))

    def sync(self):
        _check(lib().simr_sync(self.h))

    def stream(self):
        return lib().simr_stream(self.h)

    # -- outputs
    def result(self, out=None):
        """The output volume; into `out` (e.g. a pinned_empty array) when given."""
        if out is None:
            out = np.empty(output_shape(self.p), np.float32)
        elif out.dtype != np.float32 or out.size != int(np.prod(output_shape(self.p))) or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous float32 array of the output's size")
        _check(lib().simr_get_result_host(self.h, _fp(out)))
        return out

    def result_device_ptr(self):
        return lib().simr_result_dev(self.h)

    def result_to_device(self, ptr):
        """Copy the result into caller-owned device memory (simr_get_result_dev)."""
        _check(lib().simr_get_result_dev(self.h, C.c_void_p(ptr)))

    def fit(self, d):
        f = SimrFit()
        _check(lib().simr_get_fit(self.h, d, C.byref(f)))
        return f

    def set_fit(self, d, f):
        _check(lib().simr_set_fit(self.h, d, C.byref(f)))

    def band(self, d, b):
        out = np.empty(2 * self.H, np.float32)
        _check(lib().simr_get_band_host(self.h, d, b, _fp(out)))
        return out.view(np.complex64).reshape(self.p.nz, self.p.ny, self.p.nx // 2 + 1)

    def set_band(self, d, b, arr):
        a = np.ascontiguousarray(arr, dtype=np.complex64).reshape(-1)
        _check(lib().simr_set_band_host(self.h, d, b, _fp(a.view(np.float32))))

    def filter_dir(self, d):
        _check(lib().simr_filter_dir(self.h, d))

    def debug_overlaps(self, d, o1, o2, k0x, k0y):
        nxy = self.p.nx * self.p.ny
        P = np.empty(2 * nxy, np.float64)
        sx = np.empty(nxy, np.float64)
        sy = np.empty(nxy, np.float64)
        _check(lib().simr_debug_overlaps(self.h, d, o1, o2, k0x, k0y, _dp(P), _dp(sx), _dp(sy)))
        shp = (self.p.ny, self.p.nx)
        return dict(P=P.view(np.complex128).reshape(shp), sx=sx.reshape(shp), sy=sy.reshape(shp))

    def overlap_volumes(self, d, o1, o2, k0x, k0y):
        """(overlap0, overlap1) in real space, complex64 [nz][ny][nx] each."""
        n = 2 * self.p.nx * self.p.ny * self.p.nz
        ov0 = np.empty(n, np.float32)
        ov1 = np.empty(n, np.float32)
        _check(lib().simr_debug_overlap_volumes(self.h, d, o1, o2, k0x, k0y, _fp(ov0), _fp(ov1)))
        shp = (self.p.nz, self.p.ny, self.p.nx)
        return ov0.view(np.complex64).reshape(shp), ov1.view(np.complex64).reshape(shp)

    def enable_timing(self, on=True):
        _check(lib().simr_enable_timing(self.h, int(on)))

    def time_stage(self, name=None):
        """Restrict stage timing to one stage name (None: every stage)."""
        _check(lib().simr_time_stage(self.h, None if name is None else name.encode()))

    def timings(self, cap=256):
        names = (C.c_char_p * cap)()
        ms = (C.c_float * cap)()
        n = _check(lib().simr_stage_timings(self.h, names, ms, cap))
        return [(names[i].decode(), ms[i]) for i in range(n)]


def reconstruct(params, otf, raw, device=0, it=0):
    """One-shot reconstruction of one time point; returns (volume, fits)."""
    with Reconstructor(params, otf, device) as r:
        r.set_raw(raw)
        r.load_and_rescale(it)
        r.set_time(it)
        r.process_volume()
        return r.result(), [r.fit(d) for d in range(params.ndirs)]
