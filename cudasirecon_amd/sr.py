"""ctypes binding of the drop-in SR_* ABI (include/cudasirecon.h).

Mirrors the reference's library interface (src/cudaSirecon/interface.cpp:10-70)
for Python callers: SR_new_from_shape / SR_setRaw / SR_loadAndRescaleImage /
SR_setCurTimeIdx / SR_processOneVolume / SR_getResult, the command-line
constructor and its file calls, and the parameter blocks.  Errors raise
SimrError; there is no fallback when the library is missing.
"""
import ctypes as C

import numpy as np

from .engine import SimrError, lib as _engine_lib

PATH_MAX = 4096
MAX_LIST = 8


def _path_fields(*names):
    return [(n, C.c_char * PATH_MAX) for n in names]


class SRReconParams(C.Structure):
    """ReconParams (src/cudaSirecon/cudaSireconImpl.h:104-180) as the C struct SR_ReconParams."""

    i, f = C.c_int32, C.c_float
    _fields_ = (
        [("k0startangle", f), ("linespacing", f), ("na", f), ("nimm", f),
         ("ndirs", i), ("nphases", i), ("norders_output", i), ("norders", i),
         ("bTwolens", i), ("bFastSIM", i), ("bBessel", i),
         ("BesselNA", f), ("BesselLambdaEx", f), ("deskewAngle", f),
         ("extraShift", i), ("bNoRecon", i), ("cropXYto", C.c_uint32), ("bWriteTitle", i),
         ("otfcutoff", f), ("zoomfact", f), ("z_zoom", i), ("nzPadTo", i), ("explodefact", f),
         ("bFilteroverlaps", i), ("recalcarrays", i), ("napodize", i), ("bSearchforvector", i),
         ("bUseTime0k0", i), ("apodizeoutput", i), ("apoGamma", f),
         ("bSuppress_singularities", i), ("suppression_radius", i), ("bDampenOrder0", i),
         ("bFitallphases", i), ("do_rescale", i), ("equalizez", i), ("equalizet", i), ("bNoKz0", i),
         ("wiener", f), ("wienerInr", f),
         ("nforceamp", i), ("forceamp", f * MAX_LIST), ("nk0angles", i), ("k0angles", f * MAX_LIST),
         ("nxotf", i), ("nyotf", i), ("nzotf", i), ("dzPSF", f), ("dkzotf", f), ("dkrotf", f),
         ("bRadAvgOTF", i), ("bOneOTFperAngle", i), ("bFixdrift", i), ("drift_filter_fact", f),
         ("constbkgd", f), ("bBgInExtHdr", i), ("bUsecorr", i)]
        + _path_fields("corrfiles")
        + [("readoutNoiseVar", f), ("electrons_per_bit", f), ("bMakemodel", i), ("bSaveSeparated", i)]
        + _path_fields("fileSeparated")
        + [("bSaveAlignedRaw", i)]
        + _path_fields("fileRawAligned")
        + [("bSaveOverlaps", i)]
        + _path_fields("fileOverlaps")
        + [("bTIFF", i)]
        + _path_fields("ifiles", "ofiles", "otffiles")
    )
    del i, f

    def as_dict(self):
        out = {}
        for name, typ in self._fields_:
            v = getattr(self, name)
            if name in ("forceamp", "k0angles"):
                n = self.nforceamp if name == "forceamp" else self.nk0angles
                v = [v[j] for j in range(n)]
            elif isinstance(v, bytes):
                v = v.decode()
            out[name] = v
        return out


class SRImageParams(C.Structure):
    """ImageParams (src/cudaSirecon/cudaSireconImpl.h:181-195) as the C struct SR_ImageParams."""

    _fields_ = [("nx", C.c_int32), ("nx_raw", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("nz0", C.c_int32),
                ("nwaves", C.c_int16), ("wave", C.c_int16 * 5), ("ntimes", C.c_int16), ("curTimeIdx", C.c_uint16),
                ("dxy", C.c_float), ("dz", C.c_float), ("dz_raw", C.c_float), ("inscale", C.c_float)]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_}
        d["wave"] = list(self.wave)
        return d


_sr = None


def lib():
    global _sr
    if _sr is not None:
        return _sr
    L = _engine_lib()
    vp, ip, fp = C.c_void_p, C.c_int, C.POINTER(C.c_float)
    argv_t = C.POINTER(C.c_char_p)
    sigs = {
        "SR_new_from_shape": (vp, [ip, ip, ip, C.c_char_p]),
        "SR_setRaw": (ip, [vp, fp, ip, ip, ip]),
        "SR_loadAndRescaleImage": (ip, [vp, ip, ip]),
        "SR_setCurTimeIdx": (ip, [vp, ip]),
        "SR_processOneVolume": (ip, [vp]),
        "SR_getResult": (ip, [vp, fp]),
        "SR_getReconParams": (C.POINTER(SRReconParams), [vp]),
        "SR_getImageParams": (C.POINTER(SRImageParams), [vp]),
        "SR_new_from_argv": (vp, [ip, argv_t]),
        "SR_setFile": (ip, [vp, ip, ip]),
        "SR_writeResult": (ip, [vp, ip, ip]),
        "SR_closeFiles": (ip, [vp]),
        "SR_delete": (None, [vp]),
        "SR_lastError": (C.c_char_p, []),
        "SR_getNTimes": (ip, [vp]),
        "SR_parseArgs": (ip, [ip, argv_t, C.POINTER(SRReconParams), C.POINTER(SRImageParams)]),
        "SR_parseConfig": (ip, [C.c_char_p, C.POINTER(SRReconParams), C.POINTER(SRImageParams)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _sr = L
    return L


def _err():
    return lib().SR_lastError().decode()


def _check(rc):
    if rc < 0:
        raise SimrError(rc, _err())
    return rc


def _argv(args):
    arr = (C.c_char_p * (len(args) + 1))()
    for k, a in enumerate(args):
        arr[k] = a.encode() if isinstance(a, str) else a
    arr[len(args)] = None
    return arr


def parse_args(args):
    """Parse a cudasirecon command line (args[0] is the program name) and its
    config file.  Returns (status, recon, image): status 0, 1 (--help) or 2 (--version)."""
    r, im = SRReconParams(), SRImageParams()
    rc = lib().SR_parseArgs(len(args), _argv(args), C.byref(r), C.byref(im))
    _check(rc)
    return rc, r, im


def parse_config(path):
    r, im = SRReconParams(), SRImageParams()
    _check(lib().SR_parseConfig(path.encode(), C.byref(r), C.byref(im)))
    return r, im


class SIMReconstructor:
    """One SIM_Reconstructor handle (in-memory or command-line form)."""

    def __init__(self, shape=None, config=None, argv=None):
        L = lib()
        if argv is not None:
            self._argv = _argv(argv)  # the handle keeps argv for the output titles
            h = L.SR_new_from_argv(len(argv), self._argv)
        else:
            nx, ny, nimages = shape
            h = L.SR_new_from_shape(int(nx), int(ny), int(nimages), (config or "").encode())
        if not h:
            raise SimrError(-1, _err())
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().SR_delete(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def recon_params(self):
        return lib().SR_getReconParams(self._h).contents

    @property
    def image_params(self):
        return lib().SR_getImageParams(self._h).contents

    @property
    def ntimes(self):
        return _check(lib().SR_getNTimes(self._h))

    def output_shape(self):
        r, im = self.recon_params, self.image_params
        return (r.z_zoom * im.nz0, int(round(r.zoomfact * im.ny)), int(round(r.zoomfact * im.nx)))

    def set_raw(self, raw):
        a = np.ascontiguousarray(raw, dtype=np.float32)
        nz, ny, nx = a.shape
        _check(lib().SR_setRaw(self._h, a.ctypes.data_as(C.POINTER(C.c_float)), nx, ny, nz))

    def load_and_rescale_image(self, it=0, iw=0):
        _check(lib().SR_loadAndRescaleImage(self._h, it, iw))

    def set_cur_time_idx(self, it):
        _check(lib().SR_setCurTimeIdx(self._h, it))

    def process_one_volume(self):
        """True when a volume was reconstructed; False when a debug dump
        (saveprefiltered / savealignedraw / saveoverlaps) stopped processing."""
        return _check(lib().SR_processOneVolume(self._h)) == 0

    def get_result(self):
        out = np.empty(self.output_shape(), np.float32)
        _check(lib().SR_getResult(self._h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_file(self, it, iw=0):
        _check(lib().SR_setFile(self._h, it, iw))

    def write_result(self, it, iw=0):
        _check(lib().SR_writeResult(self._h, it, iw))

    def close_files(self):
        _check(lib().SR_closeFiles(self._h))

    def reconstruct(self, raw, it=0):
        """setRaw -> loadAndRescaleImage -> setCurTimeIdx -> processOneVolume -> getResult."""
        self.set_raw(raw)
        self.load_and_rescale_image(it, 0)
        self.set_cur_time_idx(it)
        self.process_one_volume()
        return self.get_result()
