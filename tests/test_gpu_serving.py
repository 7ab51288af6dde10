"""GPU checks of the serving loop (csrc/serve.cpp through include/simr.h).

* every result of a mixed stream equals the same stack reconstructed alone on
  a fresh context (bit for bit: same kernels, same inputs, no shared state);
* a time series spread over several contexts equals one sequential context
  (the time-0 k0 and bleach reference travel with simr_*_time0_state);
* a tight HBM budget evicts idle contexts and still completes every request;
* device-resident input and output give the host path's result.
"""
import sys

import numpy as np
import pytest

from cases import synthetic
from cudasirecon_amd import engine, serving

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def alone(p, otf, raw, t=0, ntimes=None):
    q = p.copy()
    if ntimes is not None:
        q.ntimes = ntimes
    with engine.Reconstructor(q, otf) as r:
        r.set_raw(raw)
        r.load_and_rescale(t)
        r.set_time(t)
        r.process_volume()
        return r.result(), [r.fit(d) for d in range(q.ndirs)]


def cases_3d_2d():
    p3, otf3, raw3a, _ = synthetic(64, 64, 8, seed=31)
    _, _, raw3b, _ = synthetic(128, 128, 8, seed=32)
    p2, otf2, raw2a, _ = synthetic(64, 64, 1, nphases=3, seed=33)
    _, _, raw2b, _ = synthetic(128, 128, 1, nphases=3, seed=34)
    return (p3, otf3, [((64, 64, 8), raw3a), ((128, 128, 8), raw3b)]), \
           (p2, otf2, [((64, 64, 1), raw2a), ((128, 128, 1), raw2b)])


def shaped(p, shape):
    q = p.copy()
    q.nx, q.ny, q.nz = shape
    return q


def test_mixed_stream_results_equal_alone():
    log("case: mixed 3D 64x64x8 / 128x128x8 and 2D 64x64 / 128x128 stream, 4 in flight")
    setups = cases_3d_2d()
    ref = {}
    for si, (p, otf, items) in enumerate(setups):
        for shape, raw in items:
            ref[(si, shape)] = alone(shaped(p, shape), otf, raw)[0]
    with serving.Server(hbm_budget=8 << 30, max_inflight=4) as srv:
        ids = [srv.add_setup(p, otf) for p, otf, _ in setups]
        jobs = []
        for rep in range(3):
            for si, (p, otf, items) in enumerate(setups):
                for shape, raw in items:
                    out = np.empty(srv.output_shape(ids[si], *shape), np.float32)
                    jobs.append((si, shape, out, srv.submit(ids[si], raw, out=out, shape=shape)))
        reused = 0
        for si, shape, out, tk in jobs:
            res = srv.wait(tk)
            reused += res.context_reused
            assert res.out_elems == out.size
            assert np.array_equal(out, ref[(si, shape)]), (si, shape)
        st = srv.stats()
        log("stats", st, "reused", reused)
        assert st["completed"] == len(jobs) and st["failed"] == 0
        assert st["contexts_created"] <= 4 * 4 and reused > 0


def test_series_over_several_contexts_equals_sequential():
    log("case: 6-frame 64x64x8 series (2%/frame bleach, two k0 draws), 3 in flight vs one context")
    p, otf, rawa, _ = synthetic(64, 64, 8, seed=41)
    _, _, rawb, _ = synthetic(64, 64, 8, seed=41, k0_jitter_px=0.05)
    T = 6
    frames = []
    for t in range(T):
        base = rawa if t % 2 == 0 else rawb
        frames.append(np.clip(np.rint(base.astype(np.float32) * 0.98 ** t), 0, 65535).astype(np.uint16))
    q = p.copy()
    q.ntimes = T
    seq, seq_fits = [], []
    with engine.Reconstructor(q, otf) as r:
        for t in range(T):
            r.set_raw(frames[t])
            r.load_and_rescale(t)
            r.set_time(t)
            r.process_volume()
            seq.append(r.result())
            seq_fits.append([r.fit(d) for d in range(q.ndirs)])
    with serving.Server(hbm_budget=8 << 30, max_inflight=3) as srv:
        sid = srv.add_setup(q, otf)
        outs = [np.empty_like(seq[0]) for _ in range(T)]
        tks = [srv.submit(sid, frames[t], out=outs[t], series=7, t=t) for t in range(T)]
        results = [srv.wait(tk) for tk in tks]
        assert srv.stats()["contexts_created"] >= 2  # the series really spread over contexts
    for t in range(T):
        assert np.array_equal(outs[t], seq[t]), t
        for d in range(q.ndirs):
            assert results[t].fits[d].k0x == seq_fits[t][d].k0x and results[t].fits[d].k0y == seq_fits[t][d].k0y
            assert list(results[t].fits[d].amp_re) == list(seq_fits[t][d].amp_re)
    # t > 0 reused the time-0 k0
    for t in range(1, T):
        assert results[t].fits[0].k0x == results[0].fits[0].k0x


def test_tight_budget_evicts_and_completes():
    log("case: budget of 1.2x the largest context, alternating shapes")
    p3, otf3, items = cases_3d_2d()[0]
    need = max(engine.footprint(shaped(p3, shape)) for shape, _ in items)
    ref = {shape: alone(shaped(p3, shape), otf3, raw)[0] for shape, raw in items}
    with serving.Server(hbm_budget=int(need * 1.2), max_inflight=2) as srv:
        sid = srv.add_setup(p3, otf3)
        jobs = []
        for rep in range(3):
            for shape, raw in items:
                out = np.empty(srv.output_shape(sid, *shape), np.float32)
                jobs.append((shape, out, srv.submit(sid, raw, out=out, shape=shape)))
        for shape, out, tk in jobs:
            srv.wait(tk)
            assert np.array_equal(out, ref[shape]), shape
        st = srv.stats()
        log("stats", st)
        assert st["contexts_evicted"] > 0 and st["bytes_peak"] <= int(need * 1.2)


def test_device_resident_input_and_output():
    import torch

    log("case: 64x64x8 stack and result in HBM")
    p, otf, raw, _ = synthetic(64, 64, 8, seed=31)
    ref = alone(p, otf, raw)[0]
    d_raw = torch.from_numpy(raw.astype(np.int32)).to("cuda").to(torch.int16)  # same bits as uint16
    d_out = torch.empty(ref.size, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    with serving.Server(hbm_budget=4 << 30, max_inflight=1) as srv:
        sid = srv.add_setup(p, otf)
        tk = srv.submit(sid, (d_raw.data_ptr(), True), out=d_out.data_ptr(), shape=(64, 64, 8))
        srv.wait(tk)
    got = d_out.cpu().numpy().reshape(ref.shape)
    assert np.array_equal(got, ref)


def test_busy_key_waits_instead_of_evicting():
    """A budget that holds one context per key: a request whose key's context
    is running waits for it rather than evicting the other key's idle context
    (which the next request of that key would then have to re-create)."""
    log("case: 64x64x8 and 128x128x8 interleaved 5:1, budget for one context of each, 4 in flight")
    p3, otf3, items = cases_3d_2d()[0]
    need = [engine.footprint(shaped(p3, shape)) for shape, _ in items]
    ref = {shape: alone(shaped(p3, shape), otf3, raw)[0] for shape, raw in items}
    order = [0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1]
    with serving.Server(hbm_budget=int(sum(need) * 1.05), max_inflight=4) as srv:
        sid = srv.add_setup(p3, otf3)
        jobs = []
        for k in order:
            shape, raw = items[k]
            out = np.empty(srv.output_shape(sid, *shape), np.float32)
            jobs.append((shape, out, srv.submit(sid, raw, out=out, shape=shape)))
        for shape, out, tk in jobs:
            srv.wait(tk)
            assert np.array_equal(out, ref[shape]), shape
        st = srv.stats()
        log("stats", st)
        assert st["completed"] == len(order) and st["contexts_evicted"] == 0 and st["contexts_created"] == 2
