"""GPU parity at the benchmark configurations' real sizes: the HIP path (through
the C ABI, and through the serving loop for the mixed stream and the series)
against the CPU oracle on the same seeded inputs.

Bars (every case): output rel-L2 <= 1e-5 against the FP32 oracle carrying the
same precision plan (FP32 FFT/filter/assembly, FP64 k0/modamp reductions),
k0 within 1e-3 px and every order's modamp within 1e-3 of the oracle's.  The
error against the FP64 oracle is printed for every case.

Cases:
  * plain 2D 256x256, 3 directions x 3 phases, wiener 0.001, zoom 2;
  * plain 3D 512x512x48, 3 directions x 5 phases (720 uint16 images), modamp 0.8;
  * a time series of 4 stacks of 256x256x32 (k0 drift +-0.05 px, 2 %/frame
    bleach) spread over the serving loop's contexts, against the oracle run
    sequentially over the time points (time-0 k0 reuse);
  * one request of each serving kind in one mixed stream: 2D 512x512 (3
    phases), 3D 256x256x24, 512x512x64 and 1024x1024x16 (5 phases).

The stacks are simulated on the GPU (torch) because the numpy simulation of
the larger ones takes tens of seconds; both sides then read the same uint16
stack from host memory.
"""
import functools
import math
import sys

import numpy as np
import pytest

from cases import rel_l2
from cudasirecon_amd import engine, serving, synth
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fit_errors(p, fa, fb):
    """(largest k0 difference in pixels of the image's frequency grid, largest modamp difference)."""
    dk = 1.0 / (p.nx * p.dxy)
    norders = engine.norders_of(p)
    k0e = max(math.hypot(a.k0x - b.k0x, a.k0y - b.k0y) / dk for a, b in zip(fa, fb))
    ame = max(abs(complex(a.amp_re[o], a.amp_im[o]) - complex(b.amp_re[o], b.amp_im[o]))
              for a, b in zip(fa, fb) for o in range(norders))
    return k0e, ame


def gpu_alone(p, otf, raw):
    with engine.Reconstructor(p, otf) as r:
        r.set_raw(raw)
        r.load_and_rescale(0)
        r.set_time(0)
        r.process_volume()
        return r.result(), [r.fit(d) for d in range(p.ndirs)]


def check(name, p, otf, raw, vol, fits, ref=None):
    """Asserts the bars against the FP32 oracle (fit reductions in FP64, the
    GPU's plan) and prints the FP64-oracle distance.  `ref` = (vol, fits) of
    the FP32 oracle when the caller already has it."""
    raw32 = raw.astype(np.float32)
    if ref is None:
        ref = ob.reconstruct(p, otf, raw32, 32, True)
    rv, rf = ref
    k0e, ame = fit_errors(p, fits, rf)
    e32 = rel_l2(vol, rv)
    v64, f64 = ob.reconstruct(p, otf, raw32, 64)
    e64 = rel_l2(vol, v64)
    k0e64, ame64 = fit_errors(p, fits, f64)
    log("  %s: rel-L2 vs FP32 oracle %.2e, vs FP64 oracle %.2e; k0 %.2e px (FP64 oracle %.2e), modamp %.2e "
        "(FP64 oracle %.2e)" % (name, e32, e64, k0e, k0e64, ame, ame64))
    assert np.isfinite(vol).all()
    assert k0e < 1e-3 and ame < 1e-3, (name, k0e, ame)
    assert e32 < 1e-5, (name, e32)


@functools.lru_cache(maxsize=None)
def dataset(nx, ny, nz, nphases=5, seed=0, **kw):
    return synth.make_dataset(nx, ny, nz, nphases=nphases, seed=seed, gpu=True, **kw)


def test_plain_2d_256():
    log("case: plain 2D 256x256, 3 dirs x 3 phases, wiener 0.001, zoom 2")
    p, otf, raw, _ = dataset(256, 256, 1, nphases=3, seed=20262201, wiener=0.001)
    vol, fits = gpu_alone(p, otf, raw)
    assert vol.shape == (1, 512, 512)
    check("2D 256x256", p, otf, raw, vol, fits)


def test_plain_3d_512x512x48():
    log("case: plain 3D 512x512x48, 3 dirs x 5 phases (720 uint16 images), modamp 0.8, dampenOrder0")
    p, otf, raw, truth = dataset(512, 512, 48, seed=20262202, modamps=(0.8, 0.4))
    assert raw.shape == (720, 512, 512) and raw.dtype == np.uint16
    vol, fits = gpu_alone(p, otf, raw)
    assert vol.shape == (48, 1024, 1024)
    check("3D 512x512x48", p, otf, raw, vol, fits)
    dk = 1.0 / (p.nx * p.dxy)
    k0t = max(math.hypot(f.k0x - k[0], f.k0y - k[1]) / dk for f, k in zip(fits, truth["k0"]))
    log("  k0 vs the simulation's truth: %.2e px" % k0t)
    assert k0t < 0.05


def test_series_256x256x32_through_serving():
    log("case: time series of 4 x 256x256x32 (k0 drift +-0.05 px, 2%/frame bleach) over the serving loop, "
        "3 in flight, vs the oracle run sequentially")
    T = 4
    frames, p, otf = [], None, None
    for t in range(T):
        q, tab, raw, _ = dataset(256, 256, 32, seed=20263202 + t, k0_jitter_px=0.05, modamps=(0.8, 0.4))
        if p is None:
            p, otf = q.copy(), tab
        frames.append(np.clip(np.rint(raw.astype(np.float32) * np.float32(0.98 ** t)), 0, 65535)
                      .astype(np.uint16))
    p.ntimes = T
    outs = [np.empty(engine.output_shape(p), np.float32) for _ in range(T)]
    with serving.Server(hbm_budget=4 * engine.footprint(p), max_inflight=3) as srv:
        sid = srv.add_setup(p, otf)
        tks = [srv.submit(sid, frames[t], out=outs[t], series=3, t=t) for t in range(T)]
        res = [srv.wait(tk) for tk in tks]
    o32 = ob.Oracle(p, otf, 32, True)
    o64 = ob.Oracle(p, otf, 64, False)
    for t in range(T):
        for o in (o32, o64):
            o.load(frames[t].astype(np.float32))
            o.rescale(t)
            o.set_time(t)
            o.process_volume()
        rf = [o32.fit(d) for d in range(p.ndirs)]
        k0e, ame = fit_errors(p, res[t].fits, rf)
        e32 = rel_l2(outs[t], o32.result())
        e64 = rel_l2(outs[t], o64.result(double=True))
        log("  t=%d: rel-L2 vs FP32 oracle %.2e, vs FP64 oracle %.2e; k0 %.2e px, modamp %.2e" % (t, e32, e64, k0e,
                                                                                                 ame))
        assert k0e < 1e-3 and ame < 1e-3, (t, k0e, ame)
        assert e32 < 1e-5, (t, e32)
    # t > 0 took the time-0 k0 (bUseTime0k0, cudaSirecon.cpp:296-331)
    for t in range(1, T):
        for d in range(p.ndirs):
            assert (res[t].fits[d].k0x, res[t].fits[d].k0y) == (res[0].fits[d].k0x, res[0].fits[d].k0y)


SERVING_KINDS = [  # (nx, ny, nz, nphases, seed)
    (512, 512, 1, 3, 20264202),
    (256, 256, 24, 5, 20264212),
    (512, 512, 64, 5, 20264213),
    (1024, 1024, 16, 5, 20264214),
]


@functools.lru_cache(maxsize=None)
def serving_mix():
    """One request of each kind submitted together (4 in flight, 3 optical setups for the 3D kinds)."""
    optics = [dict(wave=515, na=1.42), dict(wave=605, na=1.42), dict(wave=515, na=1.27)]
    reqs = []
    for i, (nx, ny, nz, nph, seed) in enumerate(SERVING_KINDS):
        kw = {} if nz == 1 else dict(optics[(i - 1) % 3], modamps=(0.8, 0.4))
        p, otf, raw, _ = dataset(nx, ny, nz, nphases=nph, seed=seed, **kw)
        reqs.append((p, otf, raw))
    budget = int(sum(engine.footprint(p) for p, _, _ in reqs) * 1.05)
    outs = []
    with serving.Server(hbm_budget=budget, max_inflight=4) as srv:
        tks = []
        for p, otf, raw in reqs:
            sid = srv.add_setup(p, otf)
            out = np.empty(engine.output_shape(p), np.float32)
            outs.append(out)
            tks.append(srv.submit(sid, raw, out=out, shape=(p.nx, p.ny, p.nz)))
        res = [srv.wait(tk) for tk in tks]
    return reqs, outs, res


@pytest.mark.parametrize("kind", range(len(SERVING_KINDS)))
def test_serving_mix_kind_matches_oracle(kind):
    nx, ny, nz, nph, _ = SERVING_KINDS[kind]
    log("case: serving mix (2D 512^2, 3D 256^2x24, 512^2x64, 1024^2x16 in flight together), kind %dx%dx%d, "
        "%d phases" % (nx, ny, nz, nph))
    reqs, outs, res = serving_mix()
    p, otf, raw = reqs[kind]
    assert res[kind].status == 0
    check("serving %dx%dx%d" % (nx, ny, nz), p, otf, raw, outs[kind], res[kind].fits)
