"""ctypes binding of the serving loop in libcudasirecon.so (include/simr.h, "Serving").

The request loop itself is native (csrc/serve.cpp): worker threads, context
cache, admission by HBM footprint and time-series sharing all run in C++.
This module only forwards requests and keeps the Python buffers alive until
their requests complete.
"""
import ctypes as C

import numpy as np

from ._abi import SimrRequest, SimrRequestInfo, SimrServerStats, norders_of, output_shape
from .engine import SimrError, _check, _fp, lib


class RequestResult:
    """What simr_server_wait reports for one request."""

    def __init__(self, info, ndirs):
        self.status = info.status
        self.error = info.error.decode(errors="replace")
        self.context_reused = bool(info.context_reused)
        self.queue_ms = (info.t_start_us - info.t_submit_us) * 1e-3
        self.run_ms = (info.t_end_us - info.t_start_us) * 1e-3
        self.latency_ms = (info.t_end_us - info.t_submit_us) * 1e-3
        self.t_submit_us, self.t_start_us, self.t_end_us = info.t_submit_us, info.t_start_us, info.t_end_us
        self.out_elems = info.out_elems
        self.fits = [info.fit[d] for d in range(ndirs)]


class Server:
    """Request loop on one GPU: up to max_inflight stacks at once within hbm_budget bytes."""

    def __init__(self, hbm_budget, max_inflight=4, device=0):
        h = C.c_void_p()
        _check(lib().simr_server_create(device, int(hbm_budget), int(max_inflight), C.byref(h)))
        self.h = h
        self._setups = []
        self._keep = {}

    def add_setup(self, params, otf):
        """Register an optical setup (parameters + OTF); returns its id."""
        otf = np.ascontiguousarray(otf, dtype=np.complex64)
        need = (params.ndirs if params.bOneOTFperAngle else 1) * norders_of(params) * \
            lib().simr_otf_size(C.byref(params))
        if otf.size < need:
            raise ValueError("OTF has %d complex entries, need %d" % (otf.size, need))
        sid = C.c_int()
        _check(lib().simr_server_add_setup(self.h, C.byref(params), _fp(otf.view(np.float32)), C.byref(sid)))
        self._setups.append(params.copy())
        return sid.value

    def output_shape(self, setup, nx, ny, nz):
        p = self._setups[setup].copy()
        p.nx, p.ny, p.nz = nx, ny, nz
        return output_shape(p)

    def submit(self, setup, raw, out=None, shape=None, series=-1, t=0):
        """Queue one stack.

        raw: numpy array (host, uint16 or float32, the setup's section order) or
        (device_ptr, is_u16) for a stack already in HBM; then shape=(nx, ny, nz)
        is required.  out: numpy float32 array of the output shape, or a device
        pointer (int), or None (the result stays in the context only).
        Returns a ticket for wait().
        """
        rq = SimrRequest()
        rq.setup, rq.series, rq.t = setup, series, t
        keep = []
        if isinstance(raw, tuple):
            ptr, u16 = raw
            rq.raw, rq.raw_u16, rq.raw_on_device = ptr, int(bool(u16)), 1
            nx, ny, nz = shape
        else:
            raw = np.ascontiguousarray(raw)
            if raw.dtype != np.uint16:
                raw = np.ascontiguousarray(raw, dtype=np.float32)
            keep.append(raw)
            rq.raw, rq.raw_u16, rq.raw_on_device = raw.ctypes.data, int(raw.dtype == np.uint16), 0
            p = self._setups[setup] if 0 <= setup < len(self._setups) else None
            if shape is not None:
                nx, ny, nz = shape
            elif raw.ndim == 3:
                nz_img, ny, nx = raw.shape
                nz = nz_img // (p.ndirs * p.nphases) if p is not None else 0
            else:
                raise ValueError("pass shape=(nx, ny, nz) for a flat raw array")
        if out is None:
            rq.out, rq.out_on_device = None, 0
        elif isinstance(out, int):
            rq.out, rq.out_on_device = out, 1
        else:
            if out.dtype != np.float32 or not out.flags.c_contiguous:
                raise ValueError("out must be a C-contiguous float32 array")
            keep.append(out)
            rq.out, rq.out_on_device = out.ctypes.data, 0
        rq.nx, rq.ny, rq.nz = nx, ny, nz
        ticket = C.c_int64()
        _check(lib().simr_server_submit(self.h, C.byref(rq), C.byref(ticket)))
        self._keep[ticket.value] = (keep, self._setups[setup].ndirs)  # submit succeeded: setup is valid
        return ticket.value

    def wait(self, ticket, check=True):
        info = SimrRequestInfo()
        rc = lib().simr_server_wait(self.h, ticket, C.byref(info))
        keep, ndirs = self._keep.pop(ticket, (None, 1))
        del keep
        res = RequestResult(info, ndirs)
        if rc < 0 and not res.error:
            res.status, res.error = rc, lib().simr_last_error().decode(errors="replace")
        if check and rc < 0:
            raise SimrError(rc, res.error)
        return res

    def stats(self):
        st = SimrServerStats()
        _check(lib().simr_server_stats_get(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in SimrServerStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            lib().simr_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def percentile(values, q):
    """q-th percentile (0-100) by linear interpolation between order statistics."""
    v = np.sort(np.asarray(values, dtype=np.float64))
    if v.size == 0:
        return float("nan")
    pos = (v.size - 1) * q / 100.0
    lo = int(np.floor(pos))
    hi = min(lo + 1, v.size - 1)
    return float(v[lo] + (v[hi] - v[lo]) * (pos - lo))
