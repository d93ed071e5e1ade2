"""Benchmark: ray-surface interactions/s of the FP64 tracer on one MI355X.

Workload (one step): trace_rays over the aspheric + achromat system of
raypier_optics_amd.workloads.asphere_doublet_c2 with 4,194,304 rays already
resident in HBM, through every generation (asphere, doublet, detector).
An interaction is a ray of any generation whose nearest-face search hit a
face (end_face_idx != 0xFFFFFFFF).

Prints one JSON line.  The roofline figure is for the kernel with the largest
share of the timed kernel time, from HIP events recorded on the library's own
stream; its algorithmic bytes are counted from the ray record sizes.  The CPU
baseline is the oracle (C restatement of the reference) in its OpenMP form
on the host threads this process has, over the whole input, with its
single-thread figure on a bounded sample beside it.
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
    """Rays of all generations whose nearest-face search hit a face (counted
    on the device)."""
    from raypier_optics_amd import metrics
    total = 0
    for g in traced:
        if g.n_rays:
            end = metrics._fields(g)[3]
            total += int((end != 0xFFFFFFFF).sum())
    return total


def host_threads():
    """Threads for the all-core CPU baseline: the cores this process may run
    on, at most OMP_NUM_THREADS when that is set (the host's share per GPU)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(n, 1)


def cpu_baseline(w, sample, all_core_sample):
    """The oracle (C restatement of the reference trace) on the host: one
    thread on the first `sample` rays, and its OpenMP form on all the host
    threads this process has on the first `all_core_sample` rays; all
    generations of each."""
    import oracle
    from raypier_optics_amd import workloads
    from raypier_optics_amd._scene import compile_scene

    def run(n, threads):
        sub = w.subset(np.arange(n))
        workloads.prepare_faces(sub.face_lists, sub.wavelengths, sub.max_length)
        tables = compile_scene(sub.face_lists)
        oracle.trace_rays(tables, sub.rays[:1000], sub.max_length, sub.wavelengths, sub.recursion_limit,
                          threads=threads)   # warm
        t0 = time.perf_counter()
        traced, _ = oracle.trace_rays(tables, sub.rays, sub.max_length, sub.wavelengths, sub.recursion_limit,
                                      threads=threads)
        dt = time.perf_counter() - t0
        n_int = int(sum(np.count_nonzero(g["end_face_idx"] != 0xFFFFFFFF) for g in traced))
        return n_int, dt

    threads = host_threads()
    n1, dt1 = run(sample, 1)
    nk, dtk = run(all_core_sample, threads)
    return {"value": nk / dtk, "unit": "ray-surface interactions/s", "cores": threads, "kind": "port",
            "sample": "first %d rays of the same workload, all generations, oracle OpenMP form on %d threads "
                      "(%d interactions, %.2f s)" % (all_core_sample, threads, nk, dtk),
            "nproc": os.cpu_count(), "threads": threads,
            "one_core": {"value": n1 / dt1, "cores": 1,
                         "sample": "first %d rays, all generations (%d interactions, %.1f s)" % (sample, n1, dt1)}}


def copy_calibration(torch, mib):
    """Three plain copies of `mib` MiB: a kernel of known traffic (reads and
    writes mib MiB each) for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def timed(torch, lib, run, steps, warmup):
    """Warm up, then time `steps` calls of run() with the library's kernel
    timing on; returns wall seconds, interactions per call and per-kernel
    event totals."""
    from raypier_optics_amd import _abi
    n_int = None
    for _ in range(warmup):
        traced = run()
        if n_int is None:
            n_int = interactions(traced)
        del traced
    torch.cuda.synchronize()
    lib.rpx_timing_enable(1)
    lib.rpx_timing_read(None, None, None)   # reset totals
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        traced = run()
        del traced
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    nk = len(_abi.TIMING_KERNELS)
    ms = (ctypes.c_double * nk)()
    launches = (ctypes.c_uint64 * nk)()
    nbytes = (ctypes.c_double * nk)()
    lib.rpx_timing_read(ms, launches, nbytes)
    lib.rpx_timing_enable(0)
    if n_int is None:
        n_int = interactions(run())
    return {"seconds": dt, "interactions": n_int, "ms": list(ms), "launches": list(launches),
            "bytes": list(nbytes)}


def roofline(r, measured):
    """Roofline object for the kernel with the largest share of event time."""
    from raypier_optics_amd import _abi
    names = _abi.TIMING_KERNELS
    ms, launches, nbytes = r["ms"], r["launches"], r["bytes"]
    k = int(np.argmax(ms))
    avg_ms = ms[k] / max(launches[k], 1)
    per_launch = nbytes[k] / max(launches[k], 1)
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": names[k], "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(names[k], measured) if measured else None,
            "avg_launch_ms": avg_ms, "algorithmic_bytes_per_launch": per_launch,
            "kernel_time_share": {names[i]: round(ms[i] / max(sum(ms), 1e-9), 4)
                                  for i in range(len(names)) if launches[i]}}


def serving_latency(n_jobs, qps, seed=2024, grad_fraction=0.3, budget_bytes=3 << 29, generations_per_step=2):
    """C5 stream: n_jobs random lens jobs (2-8 surfaces, rays log-uniform in
    [1e3, 2e6], built on the device) arriving as a Poisson process of rate
    qps into the continuous-batching server under a 1.5 GiB HBM budget (the
    largest job's buffers take 1.1 GiB); a grad_fraction of them also request
    d(spot RMS)/d(surface parameters) from the adjoint, which runs beside
    the server's steps."""
    from raypier_optics_amd import workloads
    from raypier_optics_amd.serving import TraceServer, run_poisson
    jobs = [workloads.lens_job_c5(seed + k, device_rays=True) for k in range(n_jobs)]
    grad = np.random.default_rng(seed).random(n_jobs) < grad_fraction
    server = TraceServer(hbm_budget_bytes=budget_bytes, max_jobs=512, max_rays_in_flight=64 << 20,
                         generations_per_step=generations_per_step)
    run_poisson(server, jobs[:4], qps, seed=seed, gradient=grad[:4])     # warm: compile, graph build
    held0, _ = server.stats()
    deferred0 = server.submits_deferred
    results, wall = run_poisson(server, jobs, qps, seed=seed + 1, gradient=grad)
    held, peak = server.stats()
    server.close()
    lat = np.array([r.latency for r in results]) * 1000.0
    n_int = sum(r.interactions for r in results)
    bad = sum(1 for r in results if r.status != 0)
    g = np.array([k for k in range(n_jobs) if grad[k]], dtype=int)
    out = {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "mean_ms": float(lat.mean()), "max_ms": float(lat.max()), "jobs": n_jobs, "qps": qps,
           "gradient_jobs": int(grad.sum()),
           "p50_ms_trace_only": float(np.percentile(lat[~grad], 50)) if (~grad).any() else None,
           "p50_ms_with_gradient": float(np.percentile(lat[g], 50)) if len(g) else None,
           "throughput_interactions_per_s": n_int / wall, "jobs_per_s": n_jobs / wall, "wall_s": wall,
           "rays_per_job_mean": float(np.mean([w.n_rays for w in jobs])), "failed_jobs": bad,
           "workload": "lens_job_c5: 2-8 sphere/conic/asphere surfaces, N-BK7/N-SF11, rays logU(1e3, 2e6), "
                       "Poisson arrivals, %d%% gradient jobs; latency from arrival to result (host submit and "
                       "scene compile included)" % int(round(100 * grad_fraction)),
           "generations_per_step": generations_per_step,
           "hbm_budget_bytes": budget_bytes, "jobs_held_for_budget": held - held0, "peak_budget_bytes": peak,
           "submissions_deferred_for_budget": server.submits_deferred - deferred0,
           "admission_wait_ms_p99": float(np.percentile([1000.0 * (r.t_admit - r.t_submit) for r in results], 99))}
    return out


def adjoint_timing(n_rays=1 << 21, seed=4, reps=5):
    """Forward spot RMS and adjoint gradient of an 8-surface lens job with
    2,097,152 rays: wall time per call (device-resident rays)."""
    import torch
    from raypier_optics_amd import workloads
    from raypier_optics_amd.adjoint import SequentialLens, _Device, _moments, _grad, _sigma
    w = workloads.lens_job_c5(seed, n_rays=n_rays, device_rays=True)
    lens = SequentialLens.from_workload(w)
    d = _Device(lens, w.collection())
    P = lens.params
    m = _moments(lens, d, P)
    _grad(lens, d, P, m)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        m = _moments(lens, d, P)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        g = _grad(lens, d, P, m)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fwd, bwd = (t1 - t0) / reps, (t2 - t1) / reps
    return {"rays": n_rays, "surfaces": lens.n_surfaces, "parameters": 6 * lens.n_surfaces,
            "forward_ms": 1000 * fwd, "gradient_ms": 1000 * bwd, "kept_rays": int(float(m[0])),
            "spot_rms_mm": float(_sigma(m)), "ray_surface_gradients_per_s": n_rays * lens.n_surfaces / bwd}


def michelson_c3_leg(n_rays=1 << 20, orders=(-1, 0, 1)):
    """C3: the Michelson with a Zernike mirror and a grating with orders -1,
    0, +1 (up to three children per hit), 1,048,576 rays to depth 12,
    float64 and FP32-split; interactions/s.

    Each form is traced once untimed and then timed, like the C2 leg's
    warm-up steps: the generations of one trace reserve about 190 GB, and
    the first allocation of that much device memory in a process costs up
    to a second when an earlier process on the device has just freed it."""
    import torch
    from raypier_optics_amd import workloads, _abi, _lib
    from raypier_optics_amd.tracer import trace_rays
    from raypier_optics_amd.split import trace_rays_split
    lib = _lib.load()
    nk = len(_abi.TIMING_KERNELS)
    torch.cuda.empty_cache()       # earlier legs' cached blocks
    out = {"workload": "michelson_c3: N-BK7 beamsplitter cube, flat + Zernike (Noll 4-15) PEC mirrors, "
                       "600 l/mm grating orders %s in one multi-order grating, relative thresholds 1e-4, depth 12"
                       % (tuple(orders),), "rays": n_rays}
    for name, tracer in (("fp64", trace_rays), ("fp32split", trace_rays_split)):
        w = workloads.michelson_c3(n_rays=n_rays, orders=orders)
        for timed_run in (False, True):
            rc = w.collection()
            rc._ensure_device()
            rc._device_wavelengths()
            torch.cuda.synchronize()
            if timed_run:
                lib.rpx_timing_enable(1)
                lib.rpx_timing_read(None, None, None)
            t0 = time.perf_counter()
            traced, _ = tracer(rc, w.face_lists, recursion_limit=w.recursion_limit, max_length=w.max_length)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            if not timed_run:
                del traced, rc
        ms = (ctypes.c_double * nk)()
        launches = (ctypes.c_uint64 * nk)()
        nbytes = (ctypes.c_double * nk)()
        lib.rpx_timing_read(ms, launches, nbytes)
        lib.rpx_timing_enable(0)
        n_int = interactions(traced)
        out[name] = {"seconds": dt, "interactions": n_int, "interactions_per_s": n_int / dt,
                     "rays_all_generations": int(sum(g.n_rays for g in traced)), "generations": len(traced),
                     "roofline": roofline({"ms": list(ms), "launches": list(launches), "bytes": list(nbytes)},
                                          None)}
        del traced, rc
        torch.cuda.empty_cache()
    out["fp32split_speedup"] = out["fp64"]["seconds"] / out["fp32split"]["seconds"]
    return out


def strehl_c2(n_rays, det, wavelength=0.55):
    """Strehl ratio of the C2 lens for its on-axis bundle at 0.55 um
    (n_rays on the same 6 mm disk, no field angle), float64 and FP32-split,
    at the float64 trace's best focus near the detector."""
    from raypier_optics_amd import workloads, metrics
    from raypier_optics_amd.tracer import trace_rays
    from raypier_optics_amd.split import SplitRayCollection, trace_rays_split
    w = workloads.asphere_doublet_c2(n_rays=n_rays, max_angle_deg=0.0, wl_range=(wavelength, wavelength))
    rc = w.collection()
    rc._ensure_device()
    split_in = SplitRayCollection.from_collection(rc)
    t64 = trace_rays(rc, w.face_lists, recursion_limit=w.recursion_limit, max_length=w.max_length)[0]
    t32 = trace_rays_split(split_in, w.face_lists, recursion_limit=w.recursion_limit, max_length=w.max_length)[0]
    hits, _ = metrics.detector_hits(t64, det)
    center = [float(x) for x in hits.mean(dim=0)]
    paths64, paths32 = metrics.detector_paths(t64, det), metrics.detector_paths(t32, det)
    point, s64 = metrics.best_focus_strehl(paths64, wavelength, center)
    s32 = metrics.strehl_ratio(paths32, wavelength, point)
    return {"wavelength_um": wavelength, "rays": n_rays, "detector_rays": int(paths64[0].shape[0]),
            "best_focus_mm": point, "strehl_fp64": s64, "strehl_fp32split": s32,
            "strehl_rel_err": abs(s32 - s64) / s64 if s64 > 0 else None}


FIELD_FLOPS_PER_PAIR = 172   # real FP ops of one gausslet-point term, per-gausslet terms hoisted (symbolic count in tests/test_flop_count.py), besides exp, sincos and 2 sqrt


def field_c4(n_gausslets=262144, grid=512):
    """C4: trace a focused Gaussian beam of n_gausslets gausslets through the
    asphere singlet, fit the modes, and sum the field on grid x grid points
    spanning 40 um around focus, in float64 and in the FP32-split form."""
    import torch
    from raypier_optics_amd import workloads, cfields
    from raypier_optics_amd.find_focus import find_ray_focus
    from raypier_optics_amd.tracer import trace_rays
    from raypier_optics_amd import _abi, _lib
    lib = _lib.load()
    nk = len(_abi.TIMING_KERNELS)
    gc, faces, info = workloads.gausslet_c4(n_gausslets=n_gausslets, grid=grid)
    gc._ensure_device()
    trace_rays(gc, faces, recursion_limit=3, max_length=200.0)      # warm (lengths are reset per call)
    torch.cuda.synchronize()
    lib.rpx_timing_enable(1)
    lib.rpx_timing_read(None, None, None)
    t0 = time.perf_counter()
    traced, _ = trace_rays(gc, faces, recursion_limit=3, max_length=200.0)
    torch.cuda.synchronize()
    t_trace = time.perf_counter() - t0
    ms = (ctypes.c_double * nk)()
    launches = (ctypes.c_uint64 * nk)()
    nbytes = (ctypes.c_double * nk)()
    lib.rpx_timing_read(ms, launches, nbytes)
    lib.rpx_timing_enable(0)
    gausslet_roofline = roofline({"ms": list(ms), "launches": list(launches), "bytes": list(nbytes)}, None)
    out = traced[2]
    focus = find_ray_focus(out)
    pts = workloads.field_grid(focus, grid, info["span_mm"])
    modes = cfields.gausslet_modes(out)
    wl = out.wavelengths
    cfields.sum_gaussian_modes(out, modes, wl, pts[:256], precision=0)      # warm
    cfields.sum_gaussian_modes(out, modes, wl, pts[:256], precision=1)
    res, kern = {}, {}
    for prec in (0, 1):
        torch.cuda.synchronize()
        lib.rpx_timing_enable(1)
        lib.rpx_timing_read(None, None, None)
        t0 = time.perf_counter()
        e = cfields.sum_gaussian_modes(out, modes, wl, pts, precision=prec, as_tensor=True)
        torch.cuda.synchronize()
        res[prec] = (time.perf_counter() - t0, e.cpu().numpy())
        ms = (ctypes.c_double * nk)()
        launches = (ctypes.c_uint64 * nk)()
        units = (ctypes.c_double * nk)()
        lib.rpx_timing_read(ms, launches, units)
        lib.rpx_timing_enable(0)
        slot = _abi.TIMING_KERNELS.index("k_field_sum" if prec == 0 else "k_field_sum_split")
        kern[prec] = (ms[slot] / max(launches[slot], 1), units[slot] / max(launches[slot], 1))
    pairs = out.n_rays * len(pts)
    e64, e32 = res[0][1], res[1][1]
    amp = np.abs(e64).max()
    mask = np.abs(e64) >= 0.1 * amp
    p64 = pairs / res[0][0]
    return {"workload": "gausslet_c4: collimated 1.064 um Gaussian beam (2 mm), N-BK7 asphere k=-0.5 A4=1e-5, "
                        "f=20 mm (NA 0.1); field on a %dx%d grid spanning 40 um at the base-ray focus" % (grid, grid),
            "gausslets": out.n_rays, "points": len(pts), "pairs": pairs, "gausslet_trace_ms": 1000 * t_trace,
            "fp64_s": res[0][0], "fp64_pairs_per_s": p64, "fp32split_s": res[1][0],
            "fp32split_pairs_per_s": pairs / res[1][0], "fp32split_speedup": res[0][0] / res[1][0],
            "max_abs_dE_over_max_E": float(np.abs(e32 - e64).max() / amp),
            "phase_err_rad_where_E_ge_0p1max": float(np.abs(np.angle(e32[mask] / e64[mask])).max()),
            "focus_mm": [float(x) for x in focus],
            "gausslet_trace_roofline": gausslet_roofline,
            "roofline": field_roofline("k_field_sum", kern[0], 78.6),
            "roofline_split": field_roofline("k_field_sum_split", kern[1], 157.3)}


def field_roofline(name, kern, peak):
    """Compute roofline of a field-sum pair kernel: its symbolic flops per
    launch (FIELD_FLOPS_PER_PAIR x pairs) over its event-timed launch."""
    avg_ms, pairs = kern
    achieved = pairs * FIELD_FLOPS_PER_PAIR / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else 0.0
    return {"bound": "valu", "kernel": name, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "flops_per_pair": FIELD_FLOPS_PER_PAIR, "avg_launch_ms": avg_ms,
            "pairs_per_launch": pairs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rays", type=int, default=4194304)
    ap.add_argument("--cpu-sample", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=128)
    ap.add_argument("--serve-repeat", type=int, default=0,
                    help="also run the serving stream this many more times, a fresh server each time (p50/p99 of each)")
    ap.add_argument("--field", type=int, default=1, help="run the C4 gausslet field leg (0 to skip)")
    ap.add_argument("--michelson", type=int, default=1, help="run the C3 Michelson leg (0 to skip)")
    ap.add_argument("--serve-qps", type=float, default=200.0)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run device-to-device copies of this size first (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: tracing shards by contiguous blocks of input rays and the field "
              "sum by evaluation points, but no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)

    import torch
    from raypier_optics_amd import workloads, _lib, metrics
    from raypier_optics_amd.tracer import trace_rays
    from raypier_optics_amd.split import SplitRayCollection, trace_rays_split
    lib = _lib.load()
    torch.cuda.set_device(0)
    if args.copy_calibration_mib:
        copy_calibration(torch, args.copy_calibration_mib)

    w = workloads.asphere_doublet_c2(n_rays=args.rays)
    rays = w.collection()
    rays._ensure_device()            # inputs resident in HBM before timing
    rays._device_wavelengths()
    torch.cuda.synchronize()

    def run64():
        return trace_rays(rays, w.face_lists, recursion_limit=w.recursion_limit, max_length=w.max_length)[0]

    split_in = SplitRayCollection.from_collection(rays)   # FP32-split input, resident

    def run32():
        return trace_rays_split(split_in, w.face_lists, recursion_limit=w.recursion_limit,
                                max_length=w.max_length)[0]

    r64 = timed(torch, lib, run64, args.steps, args.warmup)
    r32 = timed(torch, lib, run32, args.steps, args.warmup)

    # accuracy of the split form: spot RMS per wavelength bin at the detector
    det = sum(len(fl.faces) for fl in w.face_lists) - 1
    edges = np.linspace(0.45, 0.65, 9)
    t64, t32 = run64(), run32()
    s64, n64 = metrics.spot_rms_by_wavelength(t64, det, w.wavelengths, edges)
    s32, n32 = metrics.spot_rms_by_wavelength(t32, det, w.wavelengths, edges)
    p64, _ = metrics.detector_hits(t64, det)
    p32, _ = metrics.detector_hits(t32, det)
    hit_err = float((p64 - p32).abs().max()) if p64.shape == p32.shape else None
    rel = np.abs(s32 - s64) / s64
    del t64, t32, p64, p32
    strehl = strehl_c2(args.rays, det)

    value = r64["interactions"] * args.steps / r64["seconds"]
    ms_per_step = 1000.0 * r64["seconds"] / args.steps
    value32 = r32["interactions"] * args.steps / r32["seconds"]
    fp32split = {
        "value": value32, "unit": "ray-surface interactions/s", "ms_per_step": 1000.0 * r32["seconds"] / args.steps,
        "speedup_vs_fp64": value32 / value, "interactions_per_step": r32["interactions"],
        "spot_rms_rel_err_max": float(rel.max()),
        "spot_rms_rel_err_per_bin": [float(x) for x in rel],
        "spot_rms_fp64_mm_per_bin": [float(x) for x in s64],
        "wavelength_bin_edges_um": [float(x) for x in edges],
        "detector_hits_per_bin": [int(x) for x in n64],
        "detector_hit_max_abs_err_mm": hit_err,
        "strehl": strehl,
        "record_bytes": 112, "arithmetic": "f32 state and amplitudes; f64 root solves, normals, directions, "
                                          "optical path and phase",
        "roofline": roofline(r32, measured=args.rays),
    }
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
                   "rays": args.rays, "interactions_per_step": r64["interactions"],
                   "generations": w.recursion_limit,
                   "per_ray_wavelength_um": [0.45, 0.65]},
        "roofline": roofline(r64, measured=args.rays),
        "fp32split": fp32split,
        "latency": serving_latency(args.serve_jobs, args.serve_qps) if args.serve_jobs > 0 else None,
        # the same stream with one generation per server step, for comparison
        "latency_one_generation_per_step": {k: v for k, v in serving_latency(
            args.serve_jobs, args.serve_qps, generations_per_step=1).items()
            if k in ("p50_ms", "p99_ms", "mean_ms", "max_ms", "jobs_per_s")} if args.serve_jobs > 0 else None,
        # servers created and closed in turn in one process (buffer heaps, not
        # a memory pool per server, keep the later ones as fast as the first)
        "latency_repeats": [{k: v for k, v in serving_latency(args.serve_jobs, args.serve_qps, seed=2024 + k).items()
                             if k in ("p50_ms", "p99_ms", "mean_ms", "max_ms")} for k in range(args.serve_repeat)]
        if args.serve_jobs > 0 and args.serve_repeat > 0 else None,
        "adjoint": adjoint_timing(),
        "field": field_c4() if args.field else None,
        "michelson": michelson_c3_leg() if args.michelson else None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, min(args.cpu_sample, args.rays), args.rays)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
