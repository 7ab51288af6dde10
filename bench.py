#!/usr/bin/env python3
"""Benchmark: 3D-SIM reconstruction throughput (output voxels/s) on one MI355X.

Workload (default, the drop-in single-stack case): a 512x512x48 stack,
3 directions x 5 phases = 720 uint16 images (Angle->Z->Phase), 3-beam pattern,
RA OTF, wiener 0.001, dampenOrder0 -> 1024x1024x48 float32 output.  Synthetic,
seeded data (no datasets are reachable offline).

One step = one full reconstruction of the stack with the raw images already
resident in HBM: flat-field + bleach correction + apodisation + separation +
3D FFT, the k0 search and fit, Wiener filtering and zoomed assembly
(simr_load_and_rescale + simr_process_volume).

Prints one JSON line.  An untimed calibration pass records HIP events around
every stage kernel (on the stream each runs on) and picks the kernel with the
most device time per step; the timed steps then carry events around that
kernel's launches only, which give its average launch duration for the
roofline figure.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "output voxels/s, 3D-SIM 512²×48×15 imgs; p50/p99 latency on mixed stream"
HBM_PEAK_GBS = 8000.0
FP32_VECTOR_PEAK_TFLOPS = 157.3  # MI355X published dense FP32 vector peak
# stage -> the kernel(s) it launches, as rocprofv3 names them (length-templated)
KERNEL_OF_STAGE = {
    "section_sums": "simr::k_section_sums + simr::k_ratios",
    "rows_fwd": "simr::k_rows_fwd<{nx}, unsigned short>",
    "cols_fwd_y": "simr::k_fft2_cols<{ny}, false, ...>",
    "cols_fwd_z": "simr::k_fft2_cols<{nz}, false, ...>",
    "ov_x": "simr::k_overlaps_x<{nx}>",
    "ov_y": "simr::k_fft2_cols<{ny}, true, ...>",
    "ov_zsum": "simr::k_overlaps_zsum<false, false>",
    "ov_cols": "simr::k_ovf_cols<{ny}, true>",
    "ov_rows": "simr::k_ovf_rows<{nx}> + simr::k_zsum_totals",
    "modamp": "simr::k_modamp_fp64",
    "findk0_peak": "simr::k_cc_prep + 2D FFT + simr::k_argmax_*",
    "filter": "simr::k_filter<...> (every band of the stack)",
    "asm_z": "simr::k_asm_z_pair<{zdim}, true> + simr::k_asm_z_pair<{zdim}, false>",
    "asm_y": "simr::k_asm_y<{ydim}>",
    "asm_x": "simr::k_asm_x<{xdim}>",
}


def geometry(p):
    import math

    nx, ny, nz = p.nx, p.ny, p.nz
    norders = p.nphases // 2 + 1
    L = dict(nx=nx, ny=ny, nz=nz, nxh=nx // 2 + 1, nxy=nx * ny, nb=2 * norders - 1, norders=norders,
             nimg=nz * p.ndirs * p.nphases, xdim=int(round(p.zoomfact * nx)), ydim=int(round(p.zoomfact * ny)),
             zdim=p.z_zoom * nz, nasm=1 + p.ndirs * (norders - 1))
    L["H"] = L["nxh"] * ny * nz
    # makeoverlaps' axial support half-width (gpuFunctionsImpl.cu:529-548)
    lam = (p.wave / p.nimm) / 1e3
    dkz = 1.0 / (nz * p.dz) if nz > 1 else 1.0
    zd = int(math.ceil((1 - math.cos(math.asin(p.na / p.nimm))) / lam / dkz))
    L["zd_ov"] = (nz // 2 - 1 if zd > nz // 2 else zd) if nz > 1 else 0
    L["nzs"] = 2 * L["zd_ov"] + 1
    return L


def stage_bytes(p, L):
    """Algorithmic HBM bytes of one launch of each stage kernel: every input
    element the kernel needs read once, every output element written once
    (complex float 8 B, raw uint16 2 B, output float 4 B)."""
    H, nxy, nb, nzs = L["H"], L["nxy"], L["nb"], L["nzs"]
    nread = 3 if p.nz > 1 else 4  # overlap x pass: band 0 + the other order's re/im (2D fit of 1 vs 2: 4)
    return {
        "section_sums": L["nimg"] * nxy * 2,
        "rows_fwd": L["nimg"] * nxy * 2 + p.ndirs * nb * H * 8,
        "cols_fwd_y": 2 * p.ndirs * nb * H * 8,
        "cols_fwd_z": 2 * p.ndirs * nb * H * 8,
        # every direction's overlap pair in one launch: the bands over the
        # 2zd+1 support planes in, both overlaps' planes out
        "ov_x": p.ndirs * (nread * nzs * L["nxh"] * p.ny * 8 + 2 * nzs * nxy * 8),
        "ov_y": p.ndirs * 2 * 2 * nzs * nxy * 8,
        "ov_zsum": p.ndirs * (2 * nzs * nxy * 8 + 16 * nxy),
        # the two-pass form: formation and y transform (bands in, both
        # overlaps' planes out), then x transform and z sums (planes in, P out)
        "ov_cols": p.ndirs * (nread * nzs * L["nxh"] * p.ny * 8 + 2 * nzs * nxy * 8),
        "ov_rows": p.ndirs * (2 * nzs * nxy * 8 + 16 * nxy),
        # every direction's order-0 band (read and written) and side-band pairs
        "filter": p.ndirs * (2 + 4 * (L["norders"] - 1)) * H * 8,
        "asm_z": p.ndirs * nb * H * 8 + L["nasm"] * L["zdim"] * nxy * 8,
        "asm_y": L["nasm"] * L["zdim"] * nxy * 8 + L["nasm"] * L["zdim"] * L["ydim"] * p.nx * 8,
        "asm_x": L["nasm"] * L["zdim"] * L["ydim"] * p.nx * 8 + L["zdim"] * L["ydim"] * L["xdim"] * 4,
    }


def stage_fft_flops(p, L):
    """FFT flops of one launch of each transform stage: 5 N log2 N per complex
    transform of N points, 2.5 N log2 N per real one (the usual FFT count)."""
    import math

    def c(n):
        return 5.0 * n * math.log2(n) if n > 1 else 0.0

    nb, nzs = L["nb"], L["nzs"]
    return {
        "rows_fwd": p.ndirs * nb * p.nz * p.ny * 0.5 * c(p.nx),
        "cols_fwd_y": p.ndirs * nb * p.nz * L["nxh"] * c(p.ny),
        "cols_fwd_z": p.ndirs * nb * p.ny * L["nxh"] * c(p.nz),
        "ov_x": p.ndirs * 2 * nzs * p.ny * c(p.nx),
        "ov_y": p.ndirs * 2 * nzs * p.nx * c(p.ny),
        "ov_cols": p.ndirs * 2 * nzs * p.nx * c(p.ny),
        "ov_rows": p.ndirs * 2 * nzs * p.ny * c(p.nx),
        "asm_z": L["nasm"] * p.nx * p.ny * c(L["zdim"]),
        "asm_y": L["nasm"] * L["zdim"] * p.nx * c(L["ydim"]),
        "asm_x": L["nasm"] * L["zdim"] * L["ydim"] * c(L["xdim"]),
    }


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p, tab, raw, gpu_vol):
    """The CPU oracle on the same 512x512x48 stack: FP32 (with the FP64 fit
    reductions the GPU uses) on every host thread of the GPU's share (at most
    16), median of 3 runs; FP64 once; and one thread on the plain 2D 256x256
    stack (the bounded one-core sample).  Also the GPU volume's distance from
    both oracle volumes."""
    from oracle import binding as ob

    nproc = os.cpu_count() or 1
    threads = min(16, nproc)
    ob.set_threads(threads)
    raw32 = raw.astype(np.float32)
    times, ref = [], None
    for _ in range(3):
        t0 = time.perf_counter()
        vol, _ = ob.reconstruct(p, tab, raw32, 32, True)
        times.append(time.perf_counter() - t0)
        if ref is None:
            ref = vol
    t32 = float(np.median(times))
    e32 = float(np.linalg.norm(gpu_vol.ravel().astype(np.float64) - ref.ravel()) /
                np.linalg.norm(ref.ravel().astype(np.float64)))
    del ref
    t0 = time.perf_counter()
    vol64, _ = ob.reconstruct(p, tab, raw32, 64)
    t64 = time.perf_counter() - t0
    e64 = float(np.linalg.norm(gpu_vol.ravel().astype(np.float64) - vol64.ravel()) / np.linalg.norm(vol64.ravel()))
    del vol64
    # one thread: the plain 2D 256x256 stack (the 3D stack would take minutes on one core)
    from cudasirecon_amd import synth

    p2, t2, raw2, _ = synth.make_dataset(256, 256, 1, nphases=3, seed=20262201, wiener=0.001)
    ob.set_threads(1)
    t0 = time.perf_counter()
    v2, _ = ob.reconstruct(p2, t2, raw2.astype(np.float32), 32, True)
    t1c = time.perf_counter() - t0
    ob.set_threads(threads)
    out_vox = gpu_vol.size
    return {
        "value": round(out_vox / t32, 1),
        "unit": "output voxels/s",
        "cores": threads,
        "kind": "port",
        "sample": "FP32 CPU oracle (FP64 fit reductions, as the GPU), OpenMP on %d threads, the same 512x512x48 "
                  "stack (3 dirs x 5 phases) -> 1024x1024x48, median of 3 runs (%.2f, %.2f, %.2f s)"
                  % ((threads,) + tuple(times)),
        "host": {"nproc": nproc, "cpu": cpu_model()},
        "fp32_s_per_stack": round(t32, 3),
        "fp64_s_per_stack": round(t64, 3),
        "fp64_voxels_per_s": round(out_vox / t64, 1),
        "one_thread_2d_256": {"s_per_stack": round(t1c, 4), "voxels_per_s": round(v2.size / t1c, 1),
                              "workload": "plain 2D 256x256, 3 dirs x 3 phases -> 512x512, FP32 oracle, 1 thread"},
        "gpu_rel_l2_vs_fp32_oracle": e32,
        "gpu_rel_l2_vs_fp64_oracle": e64,
    }


def bench_2d(steps=50, warmup=5, with_cpu=True):
    """The reference's CPU-runnable case: plain 2D SIM, 256x256, 3 directions x
    3 phases, uint16 (seeded beads + filaments, Poisson noise, background 90),
    wiener 0.001, zoom 2 -> 512x512 f32, raw resident in HBM.  The OTF is the
    synthetic 2D model table (not a makeotf product).  With with_cpu, the FP32
    CPU oracle reconstructs the same stack (timed, the CPU baseline of this
    case) and the GPU result is compared with it."""
    from cudasirecon_amd import engine, synth

    p, tab, raw, _ = synth.make_dataset(256, 256, 1, nphases=3, seed=20262201, wiener=0.001)
    r = engine.Reconstructor(p, tab)
    r.set_raw(raw)

    def step():
        r.load_and_rescale(0)
        r.set_time(0)
        r.process_volume()

    for _ in range(warmup):
        step()
    r.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    r.sync()
    dt = (time.perf_counter() - t0) / steps
    vol = r.result()
    r.close()
    res = {"workload": "plain 2D SIM 256x256, 3 dirs x 3 phases uint16 -> 512x512 f32, raw resident in HBM",
           "ms_per_stack": round(dt * 1e3, 4), "voxels_per_s": round(vol.size / dt, 1)}
    if with_cpu:
        threads = min(16, os.cpu_count() or 1)
        os.environ["OMP_NUM_THREADS"] = str(threads)
        from oracle import binding as ob

        t0 = time.perf_counter()
        ref, _ = ob.reconstruct(p, tab, raw.astype(np.float32), 32, False)
        tc = time.perf_counter() - t0
        res["cpu_oracle_ms_per_stack"] = round(tc * 1e3, 2)
        res["cpu_threads"] = threads
        res["rel_l2_vs_fp32_oracle"] = float(np.linalg.norm((vol - ref).ravel()) / np.linalg.norm(ref.ravel()))
    return res


TRAFFIC_PROFILE = "profiles/c2_plain_hbm_traffic.md"


def profiled_traffic(stage):
    """HBM bytes per dispatch of a stage's kernel, looked up in the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload (scaled by the
    copy calibration stated there); None when the summary has no line for it.
    Counters cannot be read inside this process, so this is a lookup."""
    path = os.path.join(ROOT, TRAFFIC_PROFILE)
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 3 and parts[0] == "traffic_bytes_per_dispatch" and parts[1] == stage:
            return int(parts[2])
    return None


def bench_series(T=64, distinct=4, slots=4):
    """Time series of 64 stacks of 256x256x32 (3 dirs x 5 phases, uint16) -> 512x512x32 f32 each.

    `distinct` simulated frames (each with its own k0 draw, +-0.05 px) are
    cycled over the T time points, and frame t is scaled by the 2 %/frame
    bleach decay 0.98^t; every stack is staged in HBM before timing and every
    result is copied into its own device buffer.  Plain: one context, one
    stack after another, the time points after the first as captured graphs
    (also reported: the same with every kernel launched one by one and the
    modamps fitted on the host).  Batched: the serving loop with `slots` stacks in
    flight as one series (t = 0 first, then all others on any context with the
    time-0 k0).  The two outputs are compared element for element.
    """
    import torch

    from cudasirecon_amd import engine, serving, synth

    print("bench: series: generating %d distinct 256x256x32 frames" % distinct, file=sys.stderr, flush=True)
    bases = []
    for i in range(distinct):
        p, tab, raw, _ = synth.make_dataset(256, 256, 32, seed=20263202 + i, k0_jitter_px=0.05, gpu=True)
        bases.append(torch.from_numpy(raw.astype(np.int32)).to("cuda"))
    p.ntimes = T
    d_raw = [torch.clamp(torch.round(bases[t % distinct].float() * (0.98 ** t)), 0, 65535).to(torch.int32)
             .to(torch.int16) for t in range(T)]  # int16 storage, uint16 bits
    del bases
    out_elems = int(np.prod(engine.output_shape(p)))
    d_plain = torch.empty((T, out_elems), dtype=torch.float32, device="cuda")
    d_batch = torch.empty((T, out_elems), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    vox = T * out_elems

    def run_plain(use_graphs, dst):
        q = p.copy()
        q.use_graphs = use_graphs
        r = engine.Reconstructor(q, tab)

        def plain(n):
            for t in range(n):
                r.set_raw_device(d_raw[t].data_ptr(), u16=True)
                r.load_and_rescale(t)
                r.set_time(t)
                r.process_volume()
                r.result_to_device(dst[t].data_ptr())
            r.sync()

        plain(min(T, 4))  # warm-up (t > 0 captures the graph)
        t0 = time.perf_counter()
        plain(T)
        dt = time.perf_counter() - t0
        r.close()
        return dt

    # the same kernels launched one by one with the modamps on the host, then graphed
    t_eager = run_plain(0, d_batch)
    t_plain = run_plain(1, d_plain)
    graph_diff = float((d_batch - d_plain).abs().max().item())

    budget = (slots + 1) * engine.footprint(p)
    with serving.Server(hbm_budget=budget, max_inflight=slots) as srv:
        sid = srv.add_setup(p, tab)
        shape = (p.nx, p.ny, p.nz)

        def batched(series, n):
            tks = [srv.submit(sid, (d_raw[t].data_ptr(), True), out=d_batch[t].data_ptr(), shape=shape,
                              series=series, t=t) for t in range(n)]
            return [srv.wait(tk) for tk in tks]

        batched(0, min(T, 2 * slots))  # warm-up: creates and caches the contexts
        t0 = time.perf_counter()
        res = batched(1, T)
        t_batch = time.perf_counter() - t0
        st = srv.stats()
    diff = float((d_batch - d_plain).abs().max().item())
    return {
        "workload": "time series: %d x 256x256x32, 3 dirs x 5 phases uint16 -> 512x512x32 f32, %d distinct "
                    "simulated frames (k0 draws +-0.05 px) cycled, 2%%/frame bleach, stacks resident in HBM" % (T, distinct),
        "time_points": T,
        "plain_voxels_per_s": round(vox / t_plain, 1),
        "plain_ms_per_stack": round(t_plain / T * 1e3, 3),
        "plain_eager_ms_per_stack": round(t_eager / T * 1e3, 3),
        "graphed_vs_eager_max_abs_diff": graph_diff,
        "batched_voxels_per_s": round(vox / t_batch, 1),
        "batched_ms_per_stack": round(t_batch / T * 1e3, 3),
        "speedup_vs_plain": round(t_plain / t_batch, 3),
        "in_flight": slots,
        "contexts_created": int(st["contexts_created"]),
        "max_abs_diff_vs_plain": diff,
        "mean_latency_ms": round(float(np.mean([x.latency_ms for x in res])), 3),
    }


def bench_serving(n_requests=200, rate=20.0, slots=4, seed=20264202):
    """Serving stream: Poisson arrivals at `rate` req/s mixing 2D 512^2 (3 phases), 3D 256^2x24,
    512^2x64 and 1024^2x16 (5 phases), uint16 stacks from host memory, results back to host.

    Four optical setups stay registered (one 2D, three 3D that differ in
    emission wavelength and NA); the 3D requests take their setup in turn.
    One simulated stack per request kind is reused for every request of that
    kind.  Latency = completion - arrival (submission at the arrival time).
    Also reported: the same request list submitted all at once (closed loop)
    against one plain context per kind processing it sequentially.
    """
    import threading

    from cudasirecon_amd import engine, serving, synth

    print("bench: serving: generating the request kinds", file=sys.stderr, flush=True)
    kinds = []  # (setup index, params, raw)
    p2, t2, raw2, _ = synth.make_dataset(512, 512, 1, nphases=3, seed=seed, gpu=True)
    setups = [(p2, t2)]
    optics = [dict(wave=515, na=1.42), dict(wave=605, na=1.42), dict(wave=515, na=1.27)]
    for o in optics:
        p3, t3, _, _ = synth.make_dataset(64, 64, 24, seed=seed + 1, **o)
        setups.append((p3, t3))
    def pinned_copy(a):
        b = engine.pinned_empty(a.shape, a.dtype)
        b[...] = a
        return b

    kinds.append((0, (512, 512, 1), pinned_copy(raw2)))
    for j, (nx, ny, nz) in enumerate([(256, 256, 24), (512, 512, 64), (1024, 1024, 16)]):
        _, _, raw3, _ = synth.make_dataset(nx, ny, nz, seed=seed + 10 + j, gpu=True)
        kinds.append((None, (nx, ny, nz), pinned_copy(raw3)))
    rng = np.random.default_rng(seed)
    plan = []  # (arrival s, kind, setup)
    t = 0.0
    n3 = 0
    for i in range(n_requests):
        t += rng.exponential(1.0 / rate)
        k = int(rng.integers(0, 4))
        s = 0 if k == 0 else 1 + (n3 % 3)
        n3 += k != 0
        plan.append((t, k, s))
    shapes = {}
    for k, (_, shape, _) in enumerate(kinds):
        for s in ([0] if k == 0 else [1, 2, 3]):
            q = setups[s][0].copy()
            q.nx, q.ny, q.nz = shape
            shapes[(k, s)] = engine.output_shape(q)
    need = []
    for (k, s), _ in shapes.items():
        q = setups[s][0].copy()
        q.nx, q.ny, q.nz = kinds[k][1]
        need.append(engine.footprint(q))
    # one context of every (kind, setup) plus slots-1 more of the largest fit at once (the cache keeps the
    # OTFs and arenas resident); admission still bounds the stacks in flight to `slots`
    budget = int(min((sum(need) + (slots - 1) * max(need)) * 1.05, 200e9))
    vox_of = {ks: int(np.prod(v)) for ks, v in shapes.items()}

    # output buffers in page-locked memory, a pool per request kind (its setups share the shape)
    pools = {k: [engine.pinned_empty(shapes[ks], np.float32) for _ in range(slots + 1)]
             for k in range(len(kinds)) for ks in [next(x for x in shapes if x[0] == k)]}

    def run_stream(srv, ids, paced):
        free = {k: list(v) for k, v in pools.items()}
        cond = threading.Condition()
        lat, vox = [], [0]
        tickets = []
        done = threading.Event()

        def collector():
            i = 0
            while i < len(plan):
                with cond:
                    while i >= len(tickets):
                        cond.wait()
                    tk, ks, buf, arr_us = tickets[i]
                res = srv.wait(tk)
                lat.append((res.t_end_us - arr_us) * 1e-6)
                vox[0] += vox_of[ks]
                if (i + 1) % 250 == 0:
                    print("bench: serving (%s): %d of %d requests done" % ("paced" if paced else "closed loop",
                                                                          i + 1, len(plan)),
                          file=sys.stderr, flush=True)
                with cond:
                    free[ks[0]].append(buf)
                    cond.notify_all()
                i += 1
            done.set()

        th = threading.Thread(target=collector)
        th.start()
        t0_us = serving_now_us()
        for t_arr, k, s in plan:
            arr_us = t0_us + (t_arr * 1e6 if paced else 0.0)  # arrival time; latency counts from here
            if paced:
                dt = (arr_us - serving_now_us()) * 1e-6
                if dt > 0:
                    time.sleep(dt)
            ks = (k, s)
            with cond:
                while not free[k]:
                    cond.wait()
                buf = free[k].pop()
            tk = srv.submit(ids[s], kinds[k][2], out=buf, shape=kinds[k][1])
            with cond:
                tickets.append((tk, ks, buf, arr_us))
                cond.notify_all()
        done.wait()
        th.join()
        wall = (serving_now_us() - t0_us) * 1e-6
        return lat, vox[0], wall

    with serving.Server(hbm_budget=budget, max_inflight=slots) as srv:
        ids = [srv.add_setup(p, tab) for p, tab in setups]
        # warm-up: one request of every (kind, setup), creates the contexts
        for (k, s) in shapes:
            out = np.empty(shapes[(k, s)], np.float32)
            srv.wait(srv.submit(ids[s], kinds[k][2], out=out, shape=kinds[k][1]))
        lat, vox, wall = run_stream(srv, ids, paced=True)
        lat_c, vox_c, wall_c = run_stream(srv, ids, paced=False)
        st = srv.stats()
    # sequential plain path over the same list: one context per (kind, setup), one request after another
    ctxs = {}
    for (k, s) in shapes:
        q = setups[s][0].copy()
        q.nx, q.ny, q.nz = kinds[k][1]
        ctxs[(k, s)] = engine.Reconstructor(q, setups[s][1])
    outs = {ks: pools[ks[0]][0] for ks in shapes}

    def plain_one(k, s):
        r = ctxs[(k, s)]
        r.set_raw(kinds[k][2])
        r.load_and_rescale(0)
        r.set_time(0)
        r.process_volume()
        engine._check(engine.lib().simr_get_result_host(r.h, engine._fp(outs[(k, s)])))

    for (k, s) in shapes:
        plain_one(k, s)
    t0 = time.perf_counter()
    for i, (_, k, s) in enumerate(plan):
        plain_one(k, s)
        if (i + 1) % 250 == 0:
            print("bench: serving (sequential plain): %d of %d requests done" % (i + 1, len(plan)),
                  file=sys.stderr, flush=True)
    wall_p = time.perf_counter() - t0
    for r in ctxs.values():
        r.close()
    ms = [x * 1e3 for x in lat]
    return {
        "workload": "serving: %d requests, Poisson arrivals at %g req/s, uniform mix of 2D 512x512 (3 phases), "
                    "3D 256x256x24, 512x512x64, 1024x1024x16 (5 phases), 4 optical setups, uint16 stacks in "
                    "page-locked host memory, results to page-locked host memory" % (n_requests, rate),
        "requests": n_requests,
        "in_flight": slots,
        "p50_latency_ms": round(serving.percentile(ms, 50), 3),
        "p99_latency_ms": round(serving.percentile(ms, 99), 3),
        "mean_latency_ms": round(float(np.mean(ms)), 3),
        "paced_voxels_per_s": round(vox / wall, 1),
        "closed_loop_voxels_per_s": round(vox_c / wall_c, 1),
        "closed_loop_p50_latency_ms": round(serving.percentile([x * 1e3 for x in lat_c], 50), 3),
        "sequential_plain_voxels_per_s": round(vox_c / wall_p, 1),
        "closed_loop_speedup_vs_plain": round(wall_p / wall_c, 3),
        "contexts_created": int(st["contexts_created"]),
        "contexts_evicted": int(st["contexts_evicted"]),
        "hbm_budget_bytes": budget,
    }


def serving_now_us():
    """The steady clock the serving loop stamps its requests with (CLOCK_MONOTONIC)."""
    return time.monotonic_ns() * 1e-3


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--fit-fp32", action="store_true", help="FP32 k0/modamp reductions (reference precision)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run a plain 1 GiB device-to-device copy kernel three times, the known-size "
                         "reference that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes")
    ap.add_argument("--no-2d", action="store_true", help="skip the plain 2D 256x256 measurement")
    ap.add_argument("--no-series", action="store_true", help="skip the time-series batching measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-stream measurement")
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="requests in the serving stream (2000: the full 100 s stream at 20 req/s)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--slots", type=int, default=4, help="stacks in flight in the batched/serving forms")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU (--gpus %d not supported): stacks are independent and would shard one "
              "process per GPU, and within a stack by SIM direction with one reduce of the output volume"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    from cudasirecon_amd import engine, synth

    if args.copy_probe:
        import torch

        src = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    print("bench: generating the synthetic 512x512x48 stack", file=sys.stderr, flush=True)
    p, tab, raw, truth = synth.make_dataset(512, 512, 48, seed=20262202, modamps=(0.8, 0.4))
    p.fit_fp64 = 0 if args.fit_fp32 else 1
    r = engine.Reconstructor(p, tab)
    r.set_raw(raw)  # resident in HBM from here on
    L = geometry(p)

    def step():
        r.load_and_rescale(0)
        r.set_time(0)
        r.process_volume()

    print("bench: warmup %d" % args.warmup, file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    r.sync()

    # calibration (untimed): every stage's device time, from HIP events on the
    # stream each stage runs on, to find the kernel that takes the most time
    cal_steps = 2
    r.enable_timing(True)
    r.time_stage(None)
    r.timings()  # clear
    for _ in range(cal_steps):
        step()
    r.sync()
    agg = {}
    for name, ms in r.timings(cap=100000):
        a = agg.setdefault(name, [0.0, 0])
        a[0] += ms
        a[1] += 1
    per_step = {k: v[0] / cal_steps for k, v in agg.items()}
    launches = {k: v[1] // cal_steps for k, v in agg.items()}
    sb = stage_bytes(p, L)
    dom = max((k for k in per_step if k in sb), key=lambda k: per_step[k])

    # timed region: events around the dominant kernel's launches only
    r.time_stage(dom)
    r.timings()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    r.sync()
    dt = time.perf_counter() - t0
    dom_ms = [ms for name, ms in r.timings(cap=100000) if name == dom]
    r.enable_timing(False)
    r.time_stage(None)
    gpu_vol = r.result()
    out_vox = gpu_vol.size
    value = args.steps * out_vox / dt
    avg_ms = float(np.mean(dom_ms))
    achieved = sb[dom] / (avg_ms * 1e-3) / 1e9

    # FFT FP32 rate of each transform kernel (calibration pass) and of the whole step
    fl = stage_fft_flops(p, L)
    fft = {}
    for k, f in fl.items():
        if k in agg:
            fft[k] = round(f / (agg[k][0] / agg[k][1] * 1e-3) / 1e9, 1)
    step_flops = sum(fl[k] * launches.get(k, 0) for k in fl)

    # PCIe-inclusive rates (not the headline value): host uint16 stack in, host float32 volume
    # out, every stack; from page-locked host buffers (simr_host_alloc), then from pageable ones
    pcie_n = 3
    raw_pin = engine.pinned_empty(raw.shape, raw.dtype)
    raw_pin[...] = raw
    out_pin = engine.pinned_empty(gpu_vol.shape, np.float32)
    r.set_raw(raw_pin)
    step()
    r.result(out_pin)
    t1 = time.perf_counter()
    for _ in range(pcie_n):
        r.set_raw(raw_pin)
        step()
        r.result(out_pin)
    pcie_dt = (time.perf_counter() - t1) / pcie_n
    t1 = time.perf_counter()
    for _ in range(pcie_n):
        r.set_raw(raw)
        step()
        host_out = r.result()
    pcie_pageable_dt = (time.perf_counter() - t1) / pcie_n
    del host_out
    # two stacks in flight: a second context on a second host thread, so one stack's
    # transfers overlap the other's reconstruction (the serving loop's arrangement)
    import threading

    r2 = engine.Reconstructor(p, tab)
    raw_pin2 = engine.pinned_empty(raw.shape, raw.dtype)
    raw_pin2[...] = raw
    out_pin2 = engine.pinned_empty(gpu_vol.shape, np.float32)

    def stream_of(rr, rp, op, n):
        for _ in range(n):
            rr.set_raw(rp)
            rr.load_and_rescale(0)
            rr.set_time(0)
            rr.process_volume()
            rr.result(op)

    stream_of(r2, raw_pin2, out_pin2, 1)
    pair_n = 4
    ths = [threading.Thread(target=stream_of, args=a) for a in ((r, raw_pin, out_pin, pair_n),
                                                                 (r2, raw_pin2, out_pin2, pair_n))]
    t1 = time.perf_counter()
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    pcie_pair_dt = (time.perf_counter() - t1) / (2 * pair_n)
    r2.close()
    del raw_pin, out_pin, raw_pin2, out_pin2
    import math

    dk = 1.0 / (p.nx * p.dxy)
    k0err = max(math.hypot(r.fit(d).k0x - truth["k0"][d][0], r.fit(d).k0y - truth["k0"][d][1]) / dk
                for d in range(p.ndirs))
    r.close()

    res = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "output voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded beads+filaments, modamps 0.8/0.4, Poisson noise, uint16)",
        "config": {
            "workload": "plain 3D-SIM single stack: 512x512x48, 3 dirs x 5 phases (720 uint16 images) "
                        "-> 1024x1024x48 f32, one stack per step, raw resident in HBM",
            "global_batch": 1,
            "fit": "fp32" if args.fit_fp32 else "fp64",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "kernel_symbol": KERNEL_OF_STAGE.get(dom, dom).format(**L),
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic(dom),
            "traffic_source": "lookup: %s (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of this kernel on this workload)"
                              % TRAFFIC_PROFILE,
            "launch_ms": round(avg_ms, 4),
            "launches_timed": len(dom_ms),
            "launches_per_step": launches.get(dom, 0),
            "algorithmic_bytes": int(sb[dom]),
            "picked_by": "largest device time per step over every stage kernel (calibration pass)",
        },
        "stage_ms_per_step": {k: round(v, 4) for k, v in sorted(per_step.items(), key=lambda kv: -kv[1])},
        "stage_launches_per_step": launches,
        "stage_hbm_frac": {k: round(sb[k] / (agg[k][0] / agg[k][1] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                           for k in agg if k in sb},
        "fft_fp32": {"unit": "GFLOP/s", "count": "5 N log2 N per complex transform of N points",
                     "per_kernel": fft, "step_gflop": round(step_flops / 1e9, 2),
                     "step_rate": round(step_flops / (dt / args.steps) / 1e9, 1),
                     "peak": FP32_VECTOR_PEAK_TFLOPS * 1e3,
                     "step_frac": round(step_flops / (dt / args.steps) / 1e12 / FP32_VECTOR_PEAK_TFLOPS, 4)},
        "k0_err_vs_truth_px": round(k0err, 5),
        "pcie_inclusive": {"value": round(out_vox / pcie_pair_dt, 1), "unit": "output voxels/s",
                           "ms_per_stack": round(pcie_pair_dt * 1e3, 3),
                           "note": "host uint16 stack copied in and the float32 volume copied out each stack, "
                                   "both in page-locked host memory (simr_host_alloc), two stacks in flight "
                                   "(two contexts, two host threads)",
                           "one_in_flight_ms_per_stack": round(pcie_dt * 1e3, 3),
                           "pageable_one_in_flight_ms_per_stack": round(pcie_pageable_dt * 1e3, 3)},
    }
    if not args.no_2d:
        res["plain_2d"] = bench_2d(with_cpu=not args.no_cpu)
        print("bench: 2D %s" % json.dumps(res["plain_2d"]), file=sys.stderr, flush=True)
    if not args.no_series:
        res["batched"] = bench_series(slots=args.slots)
        print("bench: series %s" % json.dumps(res["batched"]), file=sys.stderr, flush=True)
    if not args.no_serve:
        sv = bench_serving(n_requests=args.serve_requests, rate=args.serve_rate, slots=args.slots)
        res["serving"] = sv
        res["p50_latency_ms"], res["p99_latency_ms"] = sv["p50_latency_ms"], sv["p99_latency_ms"]
        print("bench: serving %s" % json.dumps(sv), file=sys.stderr, flush=True)
    if not args.no_cpu:
        print("bench: CPU oracle baseline", file=sys.stderr, flush=True)
        res["cpu_baseline"] = cpu_baseline(p, tab, raw, gpu_vol)
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
