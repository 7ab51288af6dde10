"""CPU checks of the serving loop's C ABI: struct layouts against the C
compiler's, and the argument / admission errors that never reach the GPU."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from cudasirecon_amd import engine, serving
from cudasirecon_amd._abi import SimrRequest, SimrRequestInfo, SimrServerStats, set_params
from cudasirecon_amd.synth import make_otf_table

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "simr.h"
#define F(T, m) printf("%s.%s %zu\n", #T, #m, offsetof(T, m));
int main(void) {
  printf("simr_request %zu\n", sizeof(simr_request));
  printf("simr_request_info %zu\n", sizeof(simr_request_info));
  printf("simr_server_stats %zu\n", sizeof(simr_server_stats));
  F(simr_request, raw) F(simr_request, out) F(simr_request, series) F(simr_request, reserved)
  F(simr_request_info, t_submit_us) F(simr_request_info, out_elems) F(simr_request_info, fit)
  F(simr_request_info, error) F(simr_server_stats, bytes_peak)
  return 0;
}
"""


def test_serving_struct_layouts_match_c(tmp_path):
    src = tmp_path / "layout.c"
    exe = tmp_path / "layout"
    src.write_text(LAYOUT_C)
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                              text=True).stdout.splitlines())
    assert int(got["simr_request"]) == C.sizeof(SimrRequest)
    assert int(got["simr_request_info"]) == C.sizeof(SimrRequestInfo)
    assert int(got["simr_server_stats"]) == C.sizeof(SimrServerStats)
    for key, val in got.items():
        if "." in key:
            T, m = key.split(".")
            cls = {"simr_request": SimrRequest, "simr_request_info": SimrRequestInfo,
                   "simr_server_stats": SimrServerStats}[T]
            assert getattr(cls, m).offset == int(val), key


def _setup_params(nx=64, ny=64, nz=8):
    p = engine.default_params()
    _, og = make_otf_table(3, 1.42, 1.515, 515, 0.08)
    set_params(p, nx=nx, ny=ny, nz=nz, dxy=0.08, dz=0.125, wave=515, na=1.42, nimm=1.515, bRadAvgOTF=1, **og)
    return p


def test_server_argument_errors():
    with pytest.raises(engine.SimrError):
        serving.Server(hbm_budget=1 << 30, max_inflight=0)
    with pytest.raises(engine.SimrError):
        serving.Server(hbm_budget=0, max_inflight=1)
    p = _setup_params()
    otf = np.zeros(3 * engine.lib().simr_otf_size(C.byref(p)), np.complex64)
    with serving.Server(hbm_budget=1 << 30, max_inflight=2) as srv:
        sid = srv.add_setup(p, otf)
        raw = np.zeros((8 * 15, 64, 64), np.uint16)
        with pytest.raises(engine.SimrError, match="unknown setup"):
            srv.submit(sid + 5, raw)
        with pytest.raises(engine.SimrError, match="time 0"):
            srv.submit(sid, raw, series=3, t=1)  # t > 0 before its t = 0
        with pytest.raises(engine.SimrError, match="negative"):
            srv.submit(sid, raw, series=3, t=-1)
        with pytest.raises(engine.SimrError, match="unknown or already collected"):
            srv.wait(12345)


def test_series_fails_when_its_time0_cannot_be_admitted():
    # t = 0 has an unsupported width (never admitted: no device work), so its
    # later time point completes with an error instead of waiting forever
    p = _setup_params()
    otf = np.zeros(3 * engine.lib().simr_otf_size(C.byref(p)), np.complex64)
    with serving.Server(hbm_budget=1 << 30, max_inflight=2) as srv:
        sid = srv.add_setup(p, otf)
        bad = np.zeros((8 * 15, 64, 67), np.uint16)
        t0 = srv.submit(sid, bad, series=0, t=0)
        r0 = srv.wait(t0, check=False)
        assert r0.status == -1 and "invalid stack shape" in r0.error
        good = np.zeros((8 * 15, 64, 64), np.uint16)
        t1 = srv.submit(sid, good, series=0, t=1)
        r1 = srv.wait(t1, check=False)
        assert r1.status == -6 and "time 0 of this series failed" in r1.error
        st = srv.stats()
        assert st["submitted"] == 2 and st["failed"] == 2 and st["contexts_created"] == 0


def test_request_larger_than_budget_is_refused():
    p = _setup_params()
    otf = np.zeros(3 * engine.lib().simr_otf_size(C.byref(p)), np.complex64)
    need = engine.footprint(p)
    assert need > 0
    with serving.Server(hbm_budget=need // 2, max_inflight=1) as srv:
        sid = srv.add_setup(p, otf)
        tk = srv.submit(sid, np.zeros((8 * 15, 64, 64), np.uint16))
        r = srv.wait(tk, check=False)
        assert r.status == -3 and "more than the budget" in r.error


def test_percentile():
    assert serving.percentile([1, 2, 3, 4], 50) == 2.5
    assert serving.percentile([5], 99) == 5.0
    assert serving.percentile(list(range(101)), 99) == 99.0


def test_otf_total_counts_orders_and_directions():
    p = _setup_params()
    L = engine.lib()
    per = L.simr_otf_size(C.byref(p))
    assert L.simr_otf_total(C.byref(p)) == 3 * per
    p.bOneOTFperAngle = 1
    assert L.simr_otf_total(C.byref(p)) == 3 * 3 * per


def test_adding_setups_while_requests_run():
    """simr_server_add_setup may be called while workers run requests: the
    request path holds its own reference to the setup it runs, so growing the
    setup table under it is safe.  Without a GPU every request fails at
    context creation (and with one, the all-zero stack has no modulation to
    fit); either way every request completes and every setup id is new."""
    import threading

    p = _setup_params()
    otf = np.zeros(3 * engine.lib().simr_otf_size(C.byref(p)), np.complex64)
    raw = np.zeros((8 * 15, 64, 64), np.uint16)
    with serving.Server(hbm_budget=1 << 30, max_inflight=4) as srv:
        first = srv.add_setup(p, otf)
        ids = []

        def adder():
            for _ in range(200):
                ids.append(srv.add_setup(p, otf))

        th = threading.Thread(target=adder)
        th.start()
        tks = [srv.submit(first, raw) for _ in range(40)]
        res = [srv.wait(tk, check=False) for tk in tks]
        th.join()
        assert ids == list(range(first + 1, first + 201))
        st = srv.stats()
        assert st["submitted"] == 40 and st["completed"] + st["failed"] == 40
        assert len(res) == 40
