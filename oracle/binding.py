"""TEST INFRASTRUCTURE ONLY - ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from cudasirecon_amd._abi import SimrFit, SimrOtfInfo, SimrOtfParams, SimrParams, norders_of, output_shape

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, ip, fp, dp = C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(SimrParams), fp, ip, ip]
        L.orc_destroy.argtypes = [vp]
        for name, args in [
            ("orc_load", [vp, fp]), ("orc_rescale", [vp, ip]), ("orc_set_time", [vp, ip]),
            ("orc_process_volume", [vp]), ("orc_preprocess_bands", [vp]), ("orc_modulation", [vp, ip]),
            ("orc_filter_assemble", [vp]), ("orc_filter_dir", [vp, ip]), ("orc_assemble", [vp]),
            ("orc_get_band", [vp, ip, ip, fp]), ("orc_set_band", [vp, ip, ip, fp]),
            ("orc_get_band_d", [vp, ip, ip, dp]),
            ("orc_get_result", [vp, fp]), ("orc_get_result_d", [vp, dp]),
            ("orc_get_fit", [vp, ip, C.POINTER(SimrFit)]), ("orc_set_fit", [vp, ip, C.POINTER(SimrFit)]),
            ("orc_get_inscale", [vp, dp]), ("orc_get_zdist", [vp, C.POINTER(C.c_int)]),
            ("orc_get_sepmatrix", [vp, dp]),
            ("orc_makeoverlaps", [vp, ip, ip, ip, C.c_float, C.c_float, dp, dp, dp, fp, fp]),
            ("orc_fft3d_f", [fp, ip, ip, ip, ip]), ("orc_fft3d_d", [dp, ip, ip, ip, ip]),
            ("orc_set_flatfield", [vp, fp, fp]),
        ]:
            getattr(L, name).restype = C.c_int
            getattr(L, name).argtypes = args
        L.orc_fitparabola.restype = C.c_float
        L.orc_fitparabola.argtypes = [C.c_float] * 3
        L.orc_fitxyparabola.restype = C.c_double
        L.orc_fitxyparabola.argtypes = [C.c_double] * 6
        L.orc_findpeak.restype = C.c_int
        L.orc_findpeak.argtypes = [fp, ip, ip, fp, fp]
        L.orc_result_elems.restype = C.c_longlong
        L.orc_result_elems.argtypes = [vp]
        L.orc_last_error.restype = C.c_char_p
        L.orc_makeotf.restype = C.c_int
        L.orc_makeotf.argtypes = [C.POINTER(SimrOtfParams), fp, ip, ip, ip, fp, fp, fp, C.POINTER(SimrOtfInfo)]
        L.orc_makeotf_error.restype = C.c_char_p
        L.orc_makeotf_i2m.restype = C.c_int
        L.orc_makeotf_i2m.argtypes = [C.POINTER(SimrOtfParams), fp, ip, ip, ip, fp, C.POINTER(SimrOtfInfo)]
        L.orc_set_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        L.orc_get_threads.argtypes = []
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def set_threads(n):
    """OpenMP threads of the oracle's loops from now on."""
    _check(lib().orc_set_threads(int(n)))


def get_threads():
    return lib().orc_get_threads()


def fitparabola(a1, a2, a3):
    return lib().orc_fitparabola(a1, a2, a3)


def fitxyparabola(x1, y1, x2, y2, x3, y3):
    return lib().orc_fitxyparabola(x1, y1, x2, y2, x3, y3)


def findpeak(arr):
    a = np.ascontiguousarray(arr, dtype=np.float32)
    px, py = C.c_float(), C.c_float()
    _check(lib().orc_findpeak(_fp(a), a.shape[1], a.shape[0], C.byref(px), C.byref(py)))
    return px.value, py.value


def fft3d(a, inverse=False):
    """Oracle FFT of a complex array (1-3 dims, last axis fastest); returns a new array."""
    a = np.asarray(a)
    dims = list(a.shape[::-1]) + [1] * (3 - a.ndim)
    if a.dtype == np.complex128:
        b = np.ascontiguousarray(a).copy()
        _check(lib().orc_fft3d_d(_dp(b.view(np.float64)), dims[0], dims[1], dims[2], int(inverse)))
    else:
        b = np.ascontiguousarray(a, dtype=np.complex64).copy()
        _check(lib().orc_fft3d_f(_fp(b.view(np.float32)), dims[0], dims[1], dims[2], int(inverse)))
    return b


class Oracle:
    """One oracle reconstruction context (precision 32 or 64)."""

    def __init__(self, params, otf, precision=32, fit64=False):
        self.p = params.copy()
        self.prec = precision
        otf = np.ascontiguousarray(otf, dtype=np.complex64)
        self._otf = otf
        self.h = lib().orc_create(C.byref(self.p), _fp(otf.view(np.float32)), precision, int(fit64))
        if not self.h:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())
        self.norders = norders_of(self.p)
        self.H = (self.p.nx // 2 + 1) * self.p.ny * self.p.nz

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def load(self, raw):
        raw = np.ascontiguousarray(raw, dtype=np.float32)
        _check(lib().orc_load(self.h, _fp(raw)))

    def set_flatfield(self, background=None, slope=None):
        """Per-pixel background and slope maps ([ny][nx] floats), or None for the constants."""
        bg = None if background is None else np.ascontiguousarray(background, dtype=np.float32)
        sl = None if slope is None else np.ascontiguousarray(slope, dtype=np.float32)
        _check(lib().orc_set_flatfield(self.h, None if bg is None else _fp(bg), None if sl is None else _fp(sl)))

    def rescale(self, it=0):
        _check(lib().orc_rescale(self.h, it))

    def set_time(self, it):
        _check(lib().orc_set_time(self.h, it))

    def process_volume(self):
        _check(lib().orc_process_volume(self.h))

    def preprocess_bands(self):
        _check(lib().orc_preprocess_bands(self.h))

    def modulation(self, d):
        _check(lib().orc_modulation(self.h, d))

    def filter_assemble(self):
        _check(lib().orc_filter_assemble(self.h))

    def assemble(self):
        """Assembly of the current bands of every direction, without filtering."""
        _check(lib().orc_assemble(self.h))

    def filter_dir(self, d):
        _check(lib().orc_filter_dir(self.h, d))

    def band(self, d, b, double=False):
        if double:
            out = np.empty(2 * self.H, np.float64)
            _check(lib().orc_get_band_d(self.h, d, b, _dp(out)))
            out = out.view(np.complex128)
        else:
            out = np.empty(2 * self.H, np.float32)
            _check(lib().orc_get_band(self.h, d, b, _fp(out)))
            out = out.view(np.complex64)
        return out.reshape(self.p.nz, self.p.ny, self.p.nx // 2 + 1)

    def set_band(self, d, b, arr):
        a = np.ascontiguousarray(arr, dtype=np.complex64).reshape(-1)
        _check(lib().orc_set_band(self.h, d, b, _fp(a.view(np.float32))))

    def result(self, double=False):
        n = lib().orc_result_elems(self.h)
        if double:
            out = np.empty(n, np.float64)
            _check(lib().orc_get_result_d(self.h, _dp(out)))
        else:
            out = np.empty(n, np.float32)
            _check(lib().orc_get_result(self.h, _fp(out)))
        return out.reshape(output_shape(self.p))

    def fit(self, d):
        f = SimrFit()
        _check(lib().orc_get_fit(self.h, d, C.byref(f)))
        return f

    def set_fit(self, d, f):
        _check(lib().orc_set_fit(self.h, d, C.byref(f)))

    def inscale(self):
        v = C.c_double()
        _check(lib().orc_get_inscale(self.h, C.byref(v)))
        return v.value

    def zdist(self):
        a = (C.c_int * 3)()
        _check(lib().orc_get_zdist(self.h, a))
        return list(a)[: self.norders]

    def sepmatrix(self):
        nb = 2 * self.norders - 1
        out = np.empty(nb * self.p.nphases, np.float64)
        _check(lib().orc_get_sepmatrix(self.h, _dp(out)))
        return out.reshape(nb, self.p.nphases)

    def makeoverlaps(self, d, o1, o2, k0x, k0y, volumes=False):
        nxy = self.p.nx * self.p.ny
        P = np.empty(2 * nxy, np.float64)
        sx = np.empty(nxy, np.float64)
        sy = np.empty(nxy, np.float64)
        ov0 = ov1 = None
        f0 = f1 = None
        if volumes:
            n = 2 * nxy * self.p.nz
            ov0 = np.empty(n, np.float32)
            ov1 = np.empty(n, np.float32)
            f0, f1 = _fp(ov0), _fp(ov1)
        _check(lib().orc_makeoverlaps(self.h, d, o1, o2, k0x, k0y, _dp(P), _dp(sx), _dp(sy), f0, f1))
        shp = (self.p.ny, self.p.nx)
        res = dict(P=P.view(np.complex128).reshape(shp), sx=sx.reshape(shp), sy=sy.reshape(shp))
        if volumes:
            vs = (self.p.nz, self.p.ny, self.p.nx)
            res["ov0"] = ov0.view(np.complex64).reshape(vs)
            res["ov1"] = ov1.view(np.complex64).reshape(vs)
        return res


def makeotf(params, psf, section_bg=None, want_band0=False):
    """Oracle makeotf (src/otf/radialft.cpp restated); same contract as engine.makeotf."""
    psf = np.ascontiguousarray(psf, dtype=np.float32)
    nsec, ny, nx = psf.shape
    nz = nsec // params.nphases
    norders = (params.nphases + 1) // 2
    nbo = 2 * norders - 1 if params.five_bands else norders
    nkr = min(nx, ny) // 2 + 1
    out = np.empty(2 * nbo * nkr * nz, np.float32)
    bg = None if section_bg is None else np.ascontiguousarray(section_bg, dtype=np.float32)
    band0 = np.empty((nz, ny, nx), np.float32) if want_band0 else None
    info = SimrOtfInfo()
    rc = lib().orc_makeotf(C.byref(params), _fp(psf), nx, ny, nsec, None if bg is None else _fp(bg), _fp(out),
                           None if band0 is None else _fp(band0), C.byref(info))
    if rc != 0:
        raise RuntimeError(lib().orc_makeotf_error().decode())
    return out.view(np.complex64).reshape(nbo, nkr, nz), info, band0


def makeotf_i2m(params, i2m):
    """Oracle of makeotf -I2M's I2M OTF; same contract as engine.makeotf_i2m."""
    i2m = np.ascontiguousarray(i2m, dtype=np.float32)
    nz, ny, nx = i2m.shape
    nkr = min(nx, ny) // 2 + 1
    out = np.empty(2 * nkr * nz, np.float32)
    info = SimrOtfInfo()
    rc = lib().orc_makeotf_i2m(C.byref(params), _fp(i2m), nx, ny, nz, _fp(out), C.byref(info))
    if rc != 0:
        raise RuntimeError(lib().orc_makeotf_error().decode())
    return out.view(np.complex64).reshape(nkr, nz), info


def reconstruct(params, otf, raw, precision=32, fit64=False):
    """One-shot oracle reconstruction of one time point.  Returns (volume, fits)."""
    o = Oracle(params, otf, precision, fit64)
    o.load(raw)
    o.rescale(0)
    o.set_time(0)
    o.process_volume()
    return o.result(double=(precision == 64)), [o.fit(d) for d in range(params.ndirs)]
