"""Benchmark: ray-surface interactions/s of the FP64 tracer on one MI355X.

Workload (one step): trace_rays over the aspheric + achromat system of
raypier_optics_amd.workloads.asphere_doublet_c2 with 4,194,304 rays already
resident in HBM, through every generation (asphere, doublet, detector).
An interaction is a ray of any generation whose nearest-face search hit a
face (end_face_idx != 0xFFFFFFFF).

Prints one JSON line.  The roofline figure is for the kernel with the largest
share of the timed kernel time, from HIP events recorded on the library's own
stream; its algorithmic bytes are counted from the ray record sizes.  The CPU
baseline is the oracle (single thread, C restatement of the reference) on a
bounded sample of the same rays.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ray-surface interactions/s on 1 GPU (FP64 vs FP32-split); p50/p99 job latency"
HBM_PEAK_GBS = 8000.0


def measured_traffic(kernel, rays):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE /
    WRITE_SIZE profile of this workload (copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_asphere_doublet_c2.json")
    try:
        with open(path) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return None
    k = prof.get("kernels", {}).get(kernel)
    if prof.get("rays") != rays or k is None:
        return None
    return k["traffic_bytes_per_launch"]


def interactions(traced):
    return int(sum(int(np.count_nonzero(g.end_face_idx != 0xFFFFFFFF)) for g in traced))


def cpu_baseline(w, sample):
    import oracle
    from raypier_optics_amd import workloads
    from raypier_optics_amd._scene import compile_scene
    sub = w.subset(np.arange(sample))
    workloads.prepare_faces(sub.face_lists, sub.wavelengths, sub.max_length)
    tables = compile_scene(sub.face_lists)
    oracle.trace_rays(tables, sub.rays[:1000], sub.max_length, sub.wavelengths, sub.recursion_limit)   # warm
    t0 = time.perf_counter()
    traced, _ = oracle.trace_rays(tables, sub.rays, sub.max_length, sub.wavelengths, sub.recursion_limit)
    dt = time.perf_counter() - t0
    n_int = int(sum(np.count_nonzero(g["end_face_idx"] != 0xFFFFFFFF) for g in traced))
    return {"value": n_int / dt, "unit": "ray-surface interactions/s", "cores": 1, "kind": "port",
            "sample": "first %d rays of the same workload, all generations (%d interactions, %.1f s)" %
                      (sample, n_int, dt)}


def copy_calibration(torch, mib):
    """Three plain copies of `mib` MiB: a kernel of known traffic (reads and
    writes mib MiB each) for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rays", type=int, default=4194304)
    ap.add_argument("--cpu-sample", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run device-to-device copies of this size first (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (see DESIGN.md on sharding)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from raypier_optics_amd import workloads, _lib, _abi
    from raypier_optics_amd.tracer import trace_rays
    lib = _lib.load()
    torch.cuda.set_device(0)
    if args.copy_calibration_mib:
        copy_calibration(torch, args.copy_calibration_mib)

    w = workloads.asphere_doublet_c2(n_rays=args.rays)
    rays = w.collection()
    rays._ensure_device()            # inputs resident in HBM before timing
    rays._device_wavelengths()
    torch.cuda.synchronize()

    n_int = None
    for _ in range(args.warmup):
        traced, _ = trace_rays(rays, w.face_lists, recursion_limit=w.recursion_limit, max_length=w.max_length)
        if n_int is None:
            n_int = interactions(traced)
        del traced
    torch.cuda.synchronize()

    lib.rpx_timing_enable(1)
    lib.rpx_timing_read(None, None, None)   # reset totals
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        traced, _ = trace_rays(rays, w.face_lists, recursion_limit=w.recursion_limit, max_length=w.max_length)
        del traced
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    KERNELS = _abi.TIMING_KERNELS
    nk = len(KERNELS)
    ms = (ctypes.c_double * nk)()
    launches = (ctypes.c_uint64 * nk)()
    nbytes = (ctypes.c_double * nk)()
    lib.rpx_timing_read(ms, launches, nbytes)
    lib.rpx_timing_enable(0)
    if n_int is None:
        traced, _ = trace_rays(rays, w.face_lists, recursion_limit=w.recursion_limit, max_length=w.max_length)
        n_int = interactions(traced)

    ms_per_step = 1000.0 * dt / args.steps
    value = n_int * args.steps / dt
    k = int(np.argmax([ms[i] for i in range(nk)]))
    avg_ms = ms[k] / max(launches[k], 1)
    bytes_per_launch = nbytes[k] / max(launches[k], 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    kernel_share = {KERNELS[i]: round(ms[i] / max(sum(ms), 1e-9), 4) for i in range(nk) if launches[i]}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "ray-surface interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "asphere_doublet_c2: asphere k=-0.6 A4..A10 + cemented N-BK7/N-SF5 doublet, "
                               "MgF2 AR, planar detector; FP64 plain trace",
                   "rays": args.rays, "interactions_per_step": n_int, "generations": w.recursion_limit,
                   "per_ray_wavelength_um": [0.45, 0.65]},
        "roofline": {"bound": "hbm", "kernel": KERNELS[k], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(KERNELS[k], args.rays),
                     "avg_launch_ms": avg_ms, "algorithmic_bytes_per_launch": bytes_per_launch,
                     "kernel_time_share": kernel_share},
        "fp32split": None,
        "latency": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, min(args.cpu_sample, args.rays))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
