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

Prints one JSON line.  Kernel-level HIP-event timings are recorded on the
context's stream over the timed steps for the roofline figure.
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
# stage -> kernel name (templated on the transform length) as rocprofv3 reports it
KERNEL_OF_STAGE = {"asm_x": "simr::k_asm_x<%d>", "asm_y": "simr::k_asm_y<%d>", "asm_z": "simr::k_asm_z<%d>",
                   "rows_fwd": "simr::k_rows_fwd<%d,", "cols_fwd_y": "simr::k_fft_cols<%d,",
                   "cols_fwd_z": "simr::k_fft_cols<%d,"}
XDIM_OF_STAGE = {"asm_x": "xdim", "asm_y": "ydim", "asm_z": "zdim", "rows_fwd": "nx", "cols_fwd_y": "ny",
                 "cols_fwd_z": "nz"}


def stage_bytes(p, L):
    """Algorithmic HBM bytes of one launch of each stage kernel (reads + writes once)."""
    nx, ny, nz = p.nx, p.ny, p.nz
    nxh = nx // 2 + 1
    H = nxh * ny * nz
    nxy = nx * ny
    norders = p.nphases // 2 + 1
    nb = 2 * norders - 1
    nimg = nz * p.ndirs * p.nphases
    xdim, ydim, zdim = L["xdim"], L["ydim"], L["zdim"]
    nasm = 1 + p.ndirs * (norders - 1)
    zd = L["zd_ov"]
    nzs = 2 * zd + 1
    return {
        "section_sums": nimg * nxy * 2,
        "rows_fwd": nimg * nxy * 2 + p.ndirs * nb * H * 8,
        "cols_fwd_y": 2 * p.ndirs * nb * H * 8,
        "cols_fwd_z": 2 * p.ndirs * nb * H * 8,
        # x: band reads over the support planes + overlap rows; y: r+w; z: read + P/SX/SY
        "overlaps": 3 * H * 8 * nzs // nz + 2 * nzs * nxy * 8 + 2 * 2 * nzs * nxy * 8 + 2 * nzs * nxy * 8 + nxy * 32,
        "asm_z": p.ndirs * nb * H * 8 + nasm * zdim * nxy * 8,
        "asm_y": nasm * zdim * nxy * 8 + nasm * zdim * ydim * nx * 8,
        "asm_x": nasm * zdim * ydim * nx * 8 + zdim * ydim * xdim * 4,
        "filter": 2 * p.ndirs * nb * H * 8,  # whole stage: every band read and written once
    }


def cpu_baseline(p, tab, raw):
    """The FP32 oracle on a bounded sample of the same workload: the same 512x512x48
    stack, reconstructed once or more until about 10 s of CPU work is done."""
    threads = min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    from oracle import binding as ob

    raw32 = raw.astype(np.float32)
    runs, total, vox = 0, 0.0, 0
    while runs < 3 and (runs == 0 or total < 10.0):
        t0 = time.perf_counter()
        vol, _ = ob.reconstruct(p, tab, raw32, 32, False)
        total += time.perf_counter() - t0
        vox += vol.size
        runs += 1
    return {
        "value": round(vox / total, 1),
        "unit": "output voxels/s",
        "cores": threads,
        "kind": "port",
        "sample": "FP32 CPU oracle (OpenMP, %d threads), the same 512x512x48 stack (3 dirs x 5 phases) "
                  "-> 1024x1024x48, %d run(s) in %.1f s" % (threads, runs, total),
    }


def profiled_traffic(kernel_prefix):
    """HBM bytes per dispatch of a kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (calibrated against a plain copy); None when absent."""
    path = os.path.join(ROOT, "profiles", "c2_plain_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_dispatch" and parts[1].startswith(kernel_prefix):
            return int(parts[2])
    return None


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
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; the multi-GPU sharding is described in DESIGN.md "
              "(--gpus %d not supported)" % args.gpus, file=sys.stderr)
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
    p, tab, raw, truth = synth.make_dataset(512, 512, 48, seed=20262202)
    p.fit_fp64 = 0 if args.fit_fp32 else 1
    r = engine.Reconstructor(p, tab)
    r.set_raw(raw)  # resident in HBM from here on
    L = dict(xdim=int(round(p.zoomfact * p.nx)), ydim=int(round(p.zoomfact * p.ny)), zdim=p.z_zoom * p.nz,
             nx=p.nx, ny=p.ny, nz=p.nz)
    # overlap support half-width (makeoverlaps), needed for the byte count
    import math
    lam = (p.wave / p.nimm) / 1e3
    dkz = 1.0 / (p.nz * p.dz)
    zd = int(math.ceil((1 - math.cos(math.asin(p.na / p.nimm))) / lam / dkz))
    L["zd_ov"] = min(zd, p.nz // 2 - 1) if zd > p.nz // 2 else zd

    def step():
        r.load_and_rescale(0)
        r.set_time(0)
        r.process_volume()

    print("bench: warmup %d" % args.warmup, file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    r.sync()
    r.enable_timing(True)
    r.timings()  # clear
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    r.sync()
    dt = time.perf_counter() - t0
    stages = r.timings(cap=100000)
    r.enable_timing(False)
    out_vox = p.z_zoom * p.nz * int(round(p.zoomfact * p.ny)) * int(round(p.zoomfact * p.nx))
    value = args.steps * out_vox / dt

    # per-stage device time (HIP events on the context stream)
    agg = {}
    for name, ms in stages:
        a = agg.setdefault(name, [0.0, 0])
        a[0] += ms
        a[1] += 1
    per_step = {k: v[0] / args.steps for k, v in agg.items()}
    sb = stage_bytes(p, L)
    launch_stages = [k for k in ("asm_x", "asm_y", "asm_z", "rows_fwd", "cols_fwd_y", "cols_fwd_z") if k in agg]
    dom = max(launch_stages, key=lambda k: per_step[k])
    avg_ms = agg[dom][0] / agg[dom][1]
    achieved = sb[dom] / (avg_ms * 1e-3) / 1e9
    # PCIe-inclusive rate (not the headline value): host uint16 stack in, host float32 volume out
    pcie_n = 3
    t1 = time.perf_counter()
    for _ in range(pcie_n):
        r.set_raw(raw)
        step()
        host_out = r.result()
    pcie_dt = (time.perf_counter() - t1) / pcie_n
    del host_out
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
        "data": "synthetic (seeded beads+filaments, Poisson noise, uint16)",
        "config": {
            "workload": "plain 3D-SIM single stack: 512x512x48, 3 dirs x 5 phases (720 uint16 images) "
                        "-> 1024x1024x48 f32, one stack per step, raw resident in HBM",
            "global_batch": 1,
            "fit": "fp32" if args.fit_fp32 else "fp64",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic(KERNEL_OF_STAGE[dom] % L[XDIM_OF_STAGE.get(dom, "xdim")]),
            "launch_ms": round(avg_ms, 4),
            "algorithmic_bytes": int(sb[dom]),
        },
        "stage_ms_per_step": {k: round(v, 4) for k, v in sorted(per_step.items(), key=lambda kv: -kv[1])},
        "k0_err_vs_truth_px": round(k0err, 5),
        "pcie_inclusive": {"value": round(out_vox / pcie_dt, 1), "unit": "output voxels/s",
                           "ms_per_stack": round(pcie_dt * 1e3, 3),
                           "note": "host uint16 stack copied in and the float32 volume copied out each stack"},
    }
    if not args.no_cpu:
        print("bench: CPU oracle baseline", file=sys.stderr, flush=True)
        res["cpu_baseline"] = cpu_baseline(p, tab, raw)
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
