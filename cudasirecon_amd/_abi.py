"""ctypes mirrors of the POD structs in include/simr.h.

The field order must match include/simr.h exactly; tests/test_abi.py checks
sizeof against the C side.
"""
import ctypes as C

SIMR_MAX_DIRS = 8
SIMR_MAX_ORDERS = 3
SIMR_MAX_PHASES = 5


class SimrParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("ndirs", C.c_int32), ("nphases", C.c_int32), ("norders_output", C.c_int32),
        ("dxy", C.c_float), ("dz", C.c_float),
        ("wave", C.c_int32),
        ("ntimes", C.c_int32),
        ("k0startangle", C.c_float), ("linespacing", C.c_float), ("na", C.c_float), ("nimm", C.c_float),
        ("nk0angles", C.c_int32),
        ("k0angles", C.c_float * SIMR_MAX_DIRS),
        ("otfcutoff", C.c_float), ("zoomfact", C.c_float),
        ("z_zoom", C.c_int32),
        ("explodefact", C.c_float),
        ("bFilteroverlaps", C.c_int32), ("recalcarrays", C.c_int32), ("napodize", C.c_int32),
        ("bSearchforvector", C.c_int32), ("bUseTime0k0", C.c_int32),
        ("apodizeoutput", C.c_int32),
        ("apoGamma", C.c_float),
        ("bSuppress_singularities", C.c_int32), ("suppression_radius", C.c_int32),
        ("bDampenOrder0", C.c_int32), ("bFitallphases", C.c_int32),
        ("do_rescale", C.c_int32), ("equalizez", C.c_int32), ("equalizet", C.c_int32), ("bNoKz0", C.c_int32),
        ("wiener", C.c_float),
        ("nforceamp", C.c_int32),
        ("forceamp", C.c_float * SIMR_MAX_ORDERS),
        ("bRadAvgOTF", C.c_int32), ("bOneOTFperAngle", C.c_int32),
        ("nxotf", C.c_int32), ("nyotf", C.c_int32), ("nzotf", C.c_int32),
        ("dkrotf", C.c_float), ("dkzotf", C.c_float),
        ("constbkgd", C.c_float),
        ("bFastSIM", C.c_int32),
        ("fit_fp64", C.c_int32),
        ("nx_raw", C.c_int32), ("deskew_factor", C.c_float), ("deskew_shift", C.c_int32),
        ("bBessel", C.c_int32), ("BesselNA", C.c_float), ("BesselLambdaEx", C.c_float),
        ("use_graphs", C.c_int32), ("reserved", C.c_int32 * 8),
    ]

    def copy(self):
        p = SimrParams()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(SimrParams))
        return p

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if hasattr(v, "__len__"):
                v = list(v)
            out[name] = v
        return out


class SimrFit(C.Structure):
    _fields_ = [
        ("k0x", C.c_float), ("k0y", C.c_float),
        ("k0guess_x", C.c_float), ("k0guess_y", C.c_float),
        ("amp_re", C.c_float * SIMR_MAX_ORDERS), ("amp_im", C.c_float * SIMR_MAX_ORDERS),
        ("corr", C.c_float * SIMR_MAX_ORDERS),
        ("n_angle_evals", C.c_int32), ("n_mag_evals", C.c_int32), ("n_reductions", C.c_int32),
    ]

    def amps(self, norders):
        return [complex(self.amp_re[i], self.amp_im[i]) for i in range(norders)]


class SimrOtfParams(C.Structure):
    """simr_otf_params: makeotf options (src/otf/radialft.cpp:86-114, 1381-1553)."""

    _fields_ = [
        ("nphases", C.c_int32), ("k0angle", C.c_float), ("linespacing", C.c_float), ("bead_diameter", C.c_float),
        ("na", C.c_float), ("nimm", C.c_float), ("wavelength", C.c_int32), ("dr", C.c_float), ("dz", C.c_float),
        ("background", C.c_float), ("napodize", C.c_int32), ("do_compen", C.c_int32), ("five_bands", C.c_int32),
        ("conjugate", C.c_int32), ("dorescale", C.c_int32), ("twolens", C.c_int32), ("coherent_bsim", C.c_int32),
        ("forced_pi_shift", C.c_int32), ("ifixz", C.c_int32), ("ifixr", C.c_int32),
        ("interpkr", C.c_int32 * 2), ("leavekz", C.c_int32 * 3), ("reserved", C.c_int32 * 8),
    ]

    def copy(self):
        p = SimrOtfParams()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(SimrOtfParams))
        return p


class SimrOtfInfo(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("norders", C.c_int32),
        ("nbands_out", C.c_int32), ("nkr", C.c_int32), ("dkr", C.c_float), ("dkz", C.c_float),
        ("xc", C.c_float), ("yc", C.c_float), ("zc", C.c_float), ("phi", C.c_float * SIMR_MAX_ORDERS),
        ("scalefactor", C.c_float), ("reserved", C.c_int32 * 8),
    ]


class SimrRequest(C.Structure):
    """simr_request: one stack for the serving loop."""

    _fields_ = [
        ("setup", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("raw", C.c_void_p), ("raw_u16", C.c_int32), ("raw_on_device", C.c_int32),
        ("out", C.c_void_p), ("out_on_device", C.c_int32), ("series", C.c_int32), ("t", C.c_int32),
        ("reserved", C.c_int32 * 4),
    ]


class SimrRequestInfo(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("context_reused", C.c_int32),
        ("t_submit_us", C.c_double), ("t_start_us", C.c_double), ("t_end_us", C.c_double),
        ("out_elems", C.c_size_t), ("fit", SimrFit * SIMR_MAX_DIRS), ("error", C.c_char * 256),
    ]


class SimrServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
        ("contexts_created", C.c_int64), ("contexts_evicted", C.c_int64),
        ("bytes_in_use", C.c_size_t), ("bytes_peak", C.c_size_t),
    ]


def otf_default_params():
    """makeotf defaults (src/otf/radialft.cpp:86-114): restated for the Python side."""
    p = SimrOtfParams()
    p.nphases, p.k0angle, p.linespacing, p.bead_diameter = 5, 1.57, 0.2, 0.12
    p.na, p.nimm, p.wavelength, p.dr, p.dz = 1.4, 1.515, 525, 0.106, 0.1
    p.background, p.napodize, p.do_compen, p.ifixz, p.ifixr = -1.0, 10, 1, 1, 0
    return p


def set_otf_params(p, **kw):
    for k, v in kw.items():
        if k in ("interpkr", "leavekz"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = int(x)
        else:
            setattr(p, k, v)
    return p


def default_params():
    """Python restatement of simr_params_default() (the cudasirecon CLI defaults).

    Mirrors SetDefaultParams (src/cudaSirecon/cudaSirecon.cpp:16-72) overlaid with the
    program-option defaults (cudaSirecon.cpp:918-1005); apodizeoutput is 2 because
    gammaApo always has a value (cudaSirecon.cpp:1040-1045).
    """
    p = SimrParams()
    p.nx = p.ny = p.nz = 0
    p.ndirs = 3
    p.nphases = 5
    p.norders_output = 0
    p.dxy = 0.1
    p.dz = 0.2
    p.wave = 530
    p.ntimes = 1
    p.k0startangle = 1.648
    p.linespacing = 0.172
    p.na = 1.2
    p.nimm = 1.33
    p.nk0angles = 0
    p.otfcutoff = 0.006
    p.zoomfact = 2.0
    p.z_zoom = 1
    p.explodefact = 1.0
    p.bFilteroverlaps = 1
    p.recalcarrays = 1
    p.napodize = 10
    p.bSearchforvector = 1
    p.bUseTime0k0 = 1
    p.apodizeoutput = 2
    p.apoGamma = 1.0
    p.bSuppress_singularities = 1
    p.suppression_radius = 10
    p.bDampenOrder0 = 0
    p.bFitallphases = 1
    p.do_rescale = 1
    p.equalizez = 0
    p.equalizet = 0
    p.bNoKz0 = 0
    p.wiener = 0.01
    p.nforceamp = 0
    p.bRadAvgOTF = 0
    p.bOneOTFperAngle = 0
    p.nxotf = p.nyotf = p.nzotf = 0
    p.dkrotf = p.dkzotf = 0.0
    p.constbkgd = 0.0
    p.bFastSIM = 0
    p.fit_fp64 = 1
    p.bBessel = 0
    p.BesselNA = 0.144
    p.BesselLambdaEx = 0.488
    p.use_graphs = 1
    return p


def set_params(p, **kw):
    for k, v in kw.items():
        if k == "k0angles":
            p.nk0angles = len(v)
            for i, a in enumerate(v):
                p.k0angles[i] = a
        elif k == "forceamp":
            p.nforceamp = len(v)
            for i, a in enumerate(v):
                p.forceamp[i] = a
        else:
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
    return p


def norders_of(p):
    return p.norders_output if p.norders_output else p.nphases // 2 + 1


def output_shape(p):
    """(z, y, x) shape of the reconstructed volume."""
    import math
    zx = int(round(p.zoomfact * p.nx))
    zy = int(round(p.zoomfact * p.ny))
    return (p.z_zoom * p.nz, zy, zx)
