"""Continuous-batching trace serving (the C-ABI server of rpx_serve.hip).

Jobs are traced with the float64 generation of trace_rays (core/tracer.py:
9-47) -- the same faces, materials and child order -- but many jobs share
each launch: every server step traces one generation of every active job,
and jobs that arrive meanwhile join at the next step when the HBM budget
allows.  A job's result is a summary of its trace: generations, interactions
and the spot of its detector hits.

    server = TraceServer()
    jid = server.submit(workload)           # workloads.Workload
    while server.busy():
        for result in server.step():
            ...
"""
import ctypes
import gc
import math
import time

import numpy as np

from . import _abi
from ._lib import load, check, require_gpu, RpxError
from ._scene import compile_scene
from .workloads import prepare_faces


class JobResult(object):
    """Outcome of one job.  spot_rms is the RMS distance of the detector hits
    from their centroid (mm); times are on the server clock (seconds)."""

    def __init__(self, r, info=None):
        self.job_id = int(r.job_id)
        self.status = int(r.status)
        self.generations = int(r.generations)
        self.interactions = int(r.interactions)
        self.children = int(r.children)
        m = [float(x) for x in r.moments]
        self.detector_hits = int(round(m[0]))
        self.moments = m[:5]
        if m[0] > 0:
            c = np.array(m[1:4]) / m[0]
            self.centroid = c
            self.spot_rms = math.sqrt(max(m[4] / m[0] - float(c @ c), 0.0))
        else:
            self.centroid = None
            self.spot_rms = float("nan")
        self.t_submit, self.t_admit, self.t_finish = float(r.t_submit), float(r.t_admit), float(r.t_finish)
        self.latency = self.t_finish - self.t_submit
        self.info = info


class TraceServer(object):
    """Owns an rpx_server_t: HBM budget in bytes, at most `max_jobs` jobs on
    the device, at most `max_rays_in_flight` rays per step, and up to
    `generations_per_step` (1 to 4) generations of each job per step;
    `graph_launch` (the default) replays each step as a HIP graph; False
    launches its copies and kernels one by one."""

    def __init__(self, hbm_budget_bytes=64 << 30, max_jobs=256, max_rays_in_flight=32 << 20,
                 generations_per_step=2, graph_launch=True):
        require_gpu()
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.rpx_server_create(int(hbm_budget_bytes), int(max_jobs), int(max_rays_in_flight), None,
                                         ctypes.byref(h)))
        self.h = h
        try:
            check(self.lib.rpx_server_set_generations_per_step(h, int(generations_per_step)))
            check(self.lib.rpx_server_set_graph_launch(h, int(bool(graph_launch))))
        except Exception:
            self.lib.rpx_server_destroy(h)
            self.h = None
            raise
        self._keep = {}         # job id -> objects the device job reads (scene, wavelengths)
        self._info = {}
        self._results = (_abi.JobResult * 64)()
        self.submits_deferred = 0   # submissions run_poisson retried because waiting inputs filled the budget

    def close(self):
        if self.h is not None and self.h.value:
            self.lib.rpx_server_destroy(self.h)
        self.h = None
        self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clock(self):
        return float(self.lib.rpx_server_clock())

    def submit(self, workload, detector_face=None, info=None):
        """Queue a workloads.Workload; returns the job id."""
        import torch
        all_faces = prepare_faces(workload.face_lists, workload.wavelengths, workload.max_length)
        for fl in workload.face_lists:
            fl.sync_transforms()
        tables = compile_scene(workload.face_lists)
        handle = tables.handle()
        rc = workload.collection()
        rc.reset_length(workload.max_length)   # as trace_rays does (core/tracer.py:20)
        dev = rc._ensure_device()
        wl = rc._device_wavelengths()
        # the rays and wavelengths were written on this thread's stream; the
        # server reads them from a stream of its own (a device-wide
        # synchronize would also wait for gradient work on other streams)
        torch.cuda.current_stream().synchronize()
        if detector_face is None:
            detector_face = len(all_faces) - 1
        s = dev.struct()
        d = _abi.JobDesc(handle.value, ctypes.addressof(s), rc.n_rays, wl.data_ptr(), len(workload.wavelengths),
                         float(workload.max_length), int(workload.recursion_limit), int(detector_face), 0)
        jid = ctypes.c_uint64(0)
        check(self.lib.rpx_server_submit(self.h, ctypes.byref(d), ctypes.byref(jid)))
        self._keep[jid.value] = (tables, wl)
        self._info[jid.value] = info
        return jid.value

    def step(self):
        """Admit, trace one batched generation, return finished JobResults."""
        out = []
        while True:
            n = ctypes.c_uint32(0)
            check(self.lib.rpx_server_step(self.h, self._results, len(self._results), ctypes.byref(n)))
            for k in range(n.value):
                r = self._results[k]
                out.append(JobResult(r, self._info.pop(int(r.job_id), None)))
                self._keep.pop(int(r.job_id), None)
            if n.value < len(self._results):
                return out

    def fail_allocations_after(self, n):
        """Test hook: the job-buffer allocation after the next n fails once
        (n < 0: none)."""
        check(self.lib.rpx_server_fail_allocations_after(self.h, int(n)))

    def counts(self):
        w, a = ctypes.c_uint32(0), ctypes.c_uint32(0)
        check(self.lib.rpx_server_counts(self.h, ctypes.byref(w), ctypes.byref(a)))
        return w.value, a.value

    def stats(self):
        """(jobs held back at admission by the HBM budget, peak bytes the
        budget counted)."""
        held, peak = ctypes.c_uint64(0), ctypes.c_uint64(0)
        check(self.lib.rpx_server_stats(self.h, ctypes.byref(held), ctypes.byref(peak)))
        return held.value, peak.value

    def busy(self):
        w, a = self.counts()
        return w + a > 0


class _GradientWorker(object):
    """Gradient jobs run on a thread and a HIP stream of their own, through
    the adjoint's torch.autograd.Function (adjoint.spot_rms): a job's
    forward and backward pass overlap the server's batched trace steps, and
    each other's host work, instead of blocking the serving loop.  Both
    streams run at the default priority."""

    def __init__(self):
        import queue
        import threading
        import torch
        self.device = torch.cuda.current_device()
        self.q_in, self.q_out = queue.Queue(), queue.Queue()
        self.stream = torch.cuda.Stream()
        self.thread = threading.Thread(target=self._run, name="rpx-gradient", daemon=True)
        self.thread.start()

    def _run(self):
        import torch
        from .adjoint import SequentialLens, spot_rms
        torch.cuda.set_device(self.device)
        with torch.cuda.stream(self.stream):
            while True:
                item = self.q_in.get()
                if item is None:
                    return
                k, w = item
                try:
                    lens = SequentialLens.from_workload(w)
                    P = torch.tensor(lens.params, dtype=torch.float64, requires_grad=True)
                    sigma = spot_rms(P, lens, w.collection())
                    sigma.backward()
                    self.q_out.put((k, float(sigma.detach()), P.grad.numpy().copy(), None))
                except Exception as e:     # reported with the job
                    self.q_out.put((k, None, None, e))

    def submit(self, k, workload):
        self.q_in.put((k, workload))

    def finished(self, timeout=None):
        """Finished gradients (k, spot RMS, gradient, error); waits up to
        `timeout` seconds for the first when given."""
        import queue
        out = []
        try:
            out.append(self.q_out.get(timeout=timeout) if timeout else self.q_out.get_nowait())
            while True:
                out.append(self.q_out.get_nowait())
        except queue.Empty:
            pass
        return out

    def close(self):
        self.q_in.put(None)
        self.thread.join()


def run_poisson(server, jobs, qps, seed=0, gradient=None):
    """Submit `jobs` (a list of Workloads) at Poisson arrival times of rate
    `qps` while stepping the server; returns the JobResults in job order and
    the wall time.  A job is submitted at the first step boundary after its
    arrival time (or, when the server refuses it because waiting inputs fill
    its HBM budget, after the next step); its latency runs from its arrival
    to its result.  Jobs k with gradient[k] true also get d(spot RMS)/d(surface
    parameters) from the adjoint (result.grad, result.grad_spot): the gradient
    starts when the job is submitted and runs beside the server's steps on a
    stream of its own; the job's result is ready once its trace and its gradient both are,
    and its latency includes both.

    The objects alive when the stream starts are frozen out of Python's
    cyclic garbage collector for its duration: a full collection of a heap
    with torch imported scans about 2e5 objects (75-90 ms), which would stall
    the step loop and every job in flight; collections during the stream
    scan only the objects made since."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / qps, len(jobs)))
    worker = _GradientWorker() if gradient is not None and any(gradient) else None
    nxt = 0
    results, ids = {}, {}
    traced = {}      # k -> JobResult of a gradient job whose gradient is not back yet
    grads = {}       # k -> (spot RMS, gradient, error, server time) of a gradient job not traced yet

    def finish(k, r):
        if worker is not None and gradient[k]:
            val, g, err, t = grads.pop(k)
            if err is not None:
                raise err
            r.grad_spot, r.grad = val, g
            r.t_finish = max(r.t_finish, t)
        r.latency = r.t_finish - (t0 + arrivals[k])
        results[k] = r

    def collect(timeout=None):
        for k, val, g, err in worker.finished(timeout):
            grads[k] = (val, g, err, server.clock())
        for k in [k for k in traced if k in grads]:
            finish(k, traced.pop(k))

    gc.collect()
    gc.freeze()
    t0 = server.clock()
    try:
        while len(results) < len(jobs):
            now = server.clock() - t0
            while nxt < len(jobs) and arrivals[nxt] <= now:
                try:
                    jid = server.submit(jobs[nxt], info=nxt)
                except RpxError as e:
                    if e.code != _abi.RPX_E_CAPACITY or not server.busy():
                        raise
                    server.submits_deferred += 1
                    break             # the budget is full of waiting inputs: submit after the next step
                ids[jid] = nxt
                if worker is not None and gradient[nxt]:
                    worker.submit(nxt, jobs[nxt])
                nxt += 1
            if worker is not None:
                collect()
            if not server.busy():
                if nxt < len(jobs):   # idle until the next arrival
                    time.sleep(max(0.0, min(arrivals[nxt] - now, 1e-3)))
                elif traced:          # only gradients outstanding
                    collect(timeout=1e-3)
                continue
            for r in server.step():
                k = ids[r.job_id]
                r.grad = r.grad_spot = None
                if worker is not None and gradient[k] and k not in grads:
                    traced[k] = r
                else:
                    finish(k, r)
    finally:
        if worker is not None:
            worker.close()
        gc.unfreeze()
    wall = server.clock() - t0
    return [results[k] for k in range(len(jobs))], wall
