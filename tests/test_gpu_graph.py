"""The graphed time points (use_graphs): a time point that takes time 0's k0
runs its modamp fit, filter and assembly as one captured HIP graph with the
modamps computed on the device.  Every replay is checked against the same
kernels launched one by one with the modamps computed on the host
(use_graphs = 0): same volume bit for bit, same modamps and correlations.
The replays are interleaved with eager work on the other context and with
host synchronisations, and a parameter change between time points makes the
graph be captured again.  The last time point is also checked against the
oracle.
"""
import sys

import numpy as np
import pytest

from cases import rel_l2, synthetic
from cudasirecon_amd import engine
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def frames_of(seed, T, **kw):
    nx, ny, nz, nph = kw.pop("shape")
    p, otf, rawa, _ = synthetic(nx, ny, nz, nphases=nph, seed=seed, **kw)
    _, _, rawb, _ = synthetic(nx, ny, nz, nphases=nph, seed=seed, k0_jitter_px=0.05, **kw)
    frames = []
    for t in range(T):
        base = rawa if t % 2 == 0 else rawb
        frames.append(np.clip(np.rint(base.astype(np.float32) * 0.98 ** t), 0, 65535).astype(np.uint16))
    return p, otf, frames


def run_pair(p, otf, frames, fit64=True, update_at=None):
    """Processes the frames on a graphed and an eager context, alternating;
    returns the per-frame (volume, fits) of both."""
    qg, qe = p.copy(), p.copy()
    qg.ntimes = qe.ntimes = len(frames)
    qg.use_graphs, qe.use_graphs = 1, 0
    qg.fit_fp64 = qe.fit_fp64 = 1 if fit64 else 0
    out = {"g": [], "e": []}
    with engine.Reconstructor(qg, otf) as g, engine.Reconstructor(qe, otf) as e:
        for t, raw in enumerate(frames):
            if update_at is not None and t == update_at:
                for r, q in ((g, qg), (e, qe)):
                    q.wiener = 0.005
                    engine._check(engine.lib().simr_update_params(r.h, engine.C.byref(q)))
            for key, r in (("g", g), ("e", e)):
                r.set_raw(raw)
                r.load_and_rescale(t)
                r.set_time(t)
                r.process_volume()
                out[key].append((r.result(), [r.fit(d) for d in range(p.ndirs)]))
    return out


def assert_same(out, ndirs, norders):
    for t, ((vg, fg), (ve, fe)) in enumerate(zip(out["g"], out["e"])):
        assert np.array_equal(vg, ve), (t, float(np.abs(vg - ve).max()))
        for d in range(ndirs):
            a, b = fg[d], fe[d]
            assert (a.k0x, a.k0y) == (b.k0x, b.k0y), (t, d)
            for o in range(norders):
                assert (a.amp_re[o], a.amp_im[o]) == (b.amp_re[o], b.amp_im[o]), (t, d, o)
                assert a.corr[o] == b.corr[o], (t, d, o)


@pytest.mark.parametrize("fit64", [True, False])
def test_graphed_series_equals_eager(fit64):
    log("case: 6-frame 64x64x8 series, graphed vs eager, fit64=%d" % fit64)
    p, otf, frames = frames_of(51, 6, shape=(64, 64, 8, 5))
    out = run_pair(p, otf, frames, fit64=fit64)
    assert_same(out, p.ndirs, 3)
    # the last frame against the oracle run sequentially over the series
    q = p.copy()
    q.ntimes = len(frames)
    o = ob.Oracle(q, otf, 32, fit64)
    for t, raw in enumerate(frames):
        o.load(raw.astype(np.float32))
        o.rescale(t)
        o.set_time(t)
        o.process_volume()
    e = rel_l2(out["g"][-1][0], o.result())
    log("  last frame vs oracle rel-L2 %.2e" % e)
    assert e < 1e-5


def test_graphed_2d_chained_forced_modamps_equal_eager():
    log("case: 5-frame 2D 128x128 series, 5 phases (chained modamps), forcemodamp, graphed vs eager")
    p, otf, frames = frames_of(52, 5, shape=(128, 128, 1, 5))
    p.nforceamp, p.forceamp[0], p.forceamp[1] = 2, 0.9, 0.9
    out = run_pair(p, otf, frames)
    assert_same(out, p.ndirs, 3)


def test_graph_recaptured_after_parameter_change():
    log("case: 6-frame 64x64x8 series, wiener changed at t=3 (the graph is captured again)")
    p, otf, frames = frames_of(53, 6, shape=(64, 64, 8, 5))
    out = run_pair(p, otf, frames, update_at=3)
    assert_same(out, p.ndirs, 3)
    assert not np.array_equal(out["g"][2][0], out["g"][3][0])
