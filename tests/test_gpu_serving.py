"""Continuous-batching server (rpx_server_*): jobs traced together in one
launch per step give the same results as each job alone and as trace_rays.

Tolerances:
  * alone vs in a mix: bit-identical (every ray is traced by the same code on
    the same inputs, children are placed by a per-job scan, and the detector
    moments are summed per job in a fixed order);
  * server vs trace_rays: generations, interactions and detector hit counts
    equal; the moments (sums of hit coordinates) within 1e-12 relative, since
    the two sum the same float64 terms in different orders.
"""
import numpy as np
import pytest

from raypier_optics_amd import workloads, metrics
from raypier_optics_amd.serving import TraceServer
from raypier_optics_amd.tracer import trace_rays

pytestmark = pytest.mark.gpu

SIZES = [3000, 70000, 1000, 20000, 150000, 5000]


def _jobs(seeds=range(6), sizes=SIZES):
    return [workloads.lens_job_c5(s, n_rays=n, device_rays=True) for s, n in zip(seeds, sizes)]


def _drain(server):
    out = []
    while server.busy():
        out += server.step()
    return out


def _alone(w):
    server = TraceServer(hbm_budget_bytes=8 << 30, max_jobs=8, max_rays_in_flight=1 << 22)
    try:
        server.submit(w)
        (r,) = _drain(server)
    finally:
        server.close()
    return r


def _same(a, b, what):
    assert a.status == b.status == 0, what
    assert (a.generations, a.interactions, a.children) == (b.generations, b.interactions, b.children), what
    assert a.moments == b.moments, (what, a.moments, b.moments)


def test_server_alone_vs_mix(gpu):
    jobs = _jobs()
    alone = [_alone(w) for w in jobs]
    server = TraceServer(hbm_budget_bytes=8 << 30, max_jobs=16, max_rays_in_flight=1 << 22)
    ids = [server.submit(w) for w in jobs]
    res = {r.job_id: r for r in _drain(server)}
    server.close()
    for k, (jid, a) in enumerate(zip(ids, alone)):
        _same(a, res[jid], "job %d" % k)
        assert a.detector_hits > 0


def test_server_generations_per_step(gpu):
    """A step traces up to four generations of each job without a host round
    trip (each after the first laid out on the device from the previous
    one's counts).  Every job's generations, interactions, children and
    detector moments are bit-identical to one generation per step, in a mix
    with a budget that lets only some jobs run at once."""
    jobs = _jobs()
    res, steps = {}, {}
    for gens in (1, 2, 4):
        server = TraceServer(hbm_budget_bytes=96 << 20, max_jobs=16, max_rays_in_flight=1 << 22,
                             generations_per_step=gens)
        ids = [server.submit(w) for w in jobs]
        out, steps[gens] = {}, 0
        while server.busy():
            out.update({r.job_id: r for r in server.step()})
            steps[gens] += 1
        server.close()
        res[gens] = [out[i] for i in ids]
    for gens in (2, 4):
        for k, (a, b) in enumerate(zip(res[1], res[gens])):
            _same(a, b, "job %d, %d generations per step" % (k, gens))
    # the first step of a job traces one generation (its input buffer holds
    # only its rays); later steps trace up to the setting
    assert steps[4] < steps[2] < steps[1] * 0.75, steps
    with pytest.raises(Exception):
        TraceServer(generations_per_step=5)


def test_server_matches_trace_rays(gpu):
    for w in _jobs(seeds=range(10, 14), sizes=[4000, 30000, 2000, 60000]):
        r = _alone(w)
        traced, faces = trace_rays(w.collection(), w.face_lists, recursion_limit=w.recursion_limit,
                                   max_length=w.max_length)
        assert r.generations == len(traced)
        n_int = sum(int(np.count_nonzero(g.end_face_idx != 0xFFFFFFFF)) for g in traced)
        assert r.interactions == n_int
        assert r.children == sum(g.n_rays for g in traced[1:]) + _last_children(w, traced)
        det = w.info["detector_face"]
        pts, _ = metrics.detector_hits(traced, det)
        p = pts.cpu().numpy()
        assert r.detector_hits == len(p)
        ref = [len(p), p[:, 0].sum(), p[:, 1].sum(), p[:, 2].sum(), (p * p).sum()]
        for q in range(5):
            assert abs(r.moments[q] - ref[q]) <= 1e-12 * max(abs(ref[q]), 1.0), (q, r.moments[q], ref[q])


def _last_children(w, traced):
    """Children of the last traced generation (trace_rays traces it but does
    not return them) -- zero unless the recursion limit cut the trace."""
    if len(traced) < w.recursion_limit:
        return 0
    from raypier_optics_amd.ctracer import _TraceContext, _generation
    ctx = _TraceContext(w.face_lists, [f for fl in w.face_lists for f in fl.faces])
    return _generation(ctx, traced[-1], w.max_length, flush=False).n_rays


def test_server_budget_admits_in_order(gpu):
    jobs = _jobs(seeds=range(20, 24), sizes=[50000] * 4)
    alone = [_alone(w) for w in jobs]
    # the four input copies (counted from submission, 4 x 50k x 188 B) plus
    # one job's children buffer (1 x 50k x 188 B): the budget first lets one
    # job in, then two once the first has released its input
    server = TraceServer(hbm_budget_bytes=50 << 20, max_jobs=8, max_rays_in_flight=1 << 22)
    ids = [server.submit(w) for w in jobs]
    res = []
    active = []
    while server.busy():
        res += server.step()
        active.append(server.counts()[1])
    server.close()
    assert active[0] == 1 and max(active) < len(jobs)
    by_id = {r.job_id: r for r in res}
    admits = [by_id[j].t_admit for j in ids]
    assert admits == sorted(admits)
    for k, (jid, a) in enumerate(zip(ids, alone)):
        _same(a, by_id[jid], "job %d" % k)


def test_server_admits_while_running(gpu):
    big, small = _jobs(seeds=[30, 31], sizes=[300000, 2000])
    ref_small = _alone(small)
    server = TraceServer(hbm_budget_bytes=8 << 30, max_jobs=8, max_rays_in_flight=1 << 22)
    jb = server.submit(big)
    done = server.step()                  # big is on the device
    js = server.submit(small)             # arrives while big is running
    done += _drain(server)
    server.close()
    res = {r.job_id: r for r in done}
    assert res[js].t_admit < res[jb].t_finish
    _same(ref_small, res[js], "small job")


def test_two_servers_at_once(gpu):
    """Two servers live together, each with its own stream and buffer cache,
    stepped in turn: every job's result is bit-identical between them."""
    jobs = _jobs(seeds=range(20, 26))
    servers = [TraceServer(hbm_budget_bytes=64 << 20, max_jobs=4, max_rays_in_flight=1 << 22) for _ in range(2)]
    try:
        ids = [[s.submit(w) for w in jobs] for s in servers]
        res = [{}, {}]
        while any(s.busy() for s in servers):
            for k, s in enumerate(servers):
                if s.busy():
                    res[k].update({r.job_id: r for r in s.step()})
    finally:
        for s in servers:
            s.close()
    for k in range(len(jobs)):
        a, b = res[0][ids[0][k]], res[1][ids[1][k]]
        _same(a, b, "job %d" % k)
        assert a.detector_hits > 0


def test_allocation_failures_end_only_their_job(gpu):
    """One job-buffer allocation made to fail (rpx_server_fail_allocations_after)
    at admission or between generations: the job that hit the failure
    finishes with RPX_E_HIP, every other job completes bit-identical to its
    run alone, the server keeps stepping, and destroy frees everything."""
    import ctypes
    from raypier_optics_amd import _abi
    jobs = _jobs(seeds=range(4), sizes=[3000, 5000, 2000, 4000])
    alone = [_alone(w) for w in jobs]
    # allocation order: 4 input copies at submit, then per step the children
    # buffers of newly admitted jobs and of every job that continues
    for after in (5, 7, 9):
        server = TraceServer(hbm_budget_bytes=8 << 30, max_jobs=16, max_rays_in_flight=1 << 22)
        ids = [server.submit(w) for w in jobs]
        server.fail_allocations_after(after - 4)
        res = {r.job_id: r for r in _drain(server)}
        w, a = server.counts()
        assert (w, a) == (0, 0)
        server.close()
        failed = [k for k, jid in enumerate(ids) if res[jid].status != 0]
        assert len(failed) == 1, (after, failed)
        for k, jid in enumerate(ids):
            r = res[jid]
            if r.status != 0:
                assert r.status == _abi.RPX_E_HIP
            else:
                _same(alone[k], r, "job %d, failure after %d allocations" % (k, after))
        print("failure after %d allocations: jobs %s failed" % (after, failed))


def test_submit_respects_budget(gpu):
    """Waiting jobs' input copies count against the HBM budget: a submit that
    would exceed it is refused with RPX_E_CAPACITY (and the server still
    runs the jobs it accepted)."""
    from raypier_optics_amd._lib import RpxError
    from raypier_optics_amd import _abi
    jobs = _jobs(seeds=range(3), sizes=[20000, 20000, 20000])
    budget = int(2.5 * 20000 * 188)
    server = TraceServer(hbm_budget_bytes=budget, max_jobs=16, max_rays_in_flight=1 << 22)
    try:
        server.submit(jobs[0])
        server.submit(jobs[1])
        with pytest.raises(RpxError) as e:
            server.submit(jobs[2])
        assert e.value.code == _abi.RPX_E_CAPACITY
        res = _drain(server)
        assert len(res) == 2 and all(r.status == 0 for r in res)
    finally:
        server.close()


def test_run_poisson_gradients_beside_the_loop(gpu):
    """Gradient jobs run through the adjoint's autograd.Function on a stream
    and thread of their own while the server steps: every job finishes, and
    each gradient is bit-identical to the same computation run alone."""
    from raypier_optics_amd.serving import run_poisson
    from raypier_optics_amd.adjoint import SequentialLens, spot_rms_and_grad
    jobs = _jobs(range(6), [3000, 20000, 1000, 5000, 8000, 2000])
    grad = [True, False, True, True, False, True]
    server = TraceServer(hbm_budget_bytes=1 << 30, max_jobs=8, max_rays_in_flight=1 << 22)
    try:
        results, wall = run_poisson(server, jobs, 2000.0, seed=1, gradient=grad)
    finally:
        server.close()
    assert len(results) == len(jobs) and wall > 0
    for k, (w, r) in enumerate(zip(jobs, results)):
        assert r.status == 0 and r.latency > 0
        if grad[k]:
            s, g = spot_rms_and_grad(SequentialLens.from_workload(w), w.collection())
            assert r.grad_spot == s and np.array_equal(r.grad, g)
        else:
            assert r.grad is None


def test_run_poisson_budget_backpressure(gpu):
    """A budget that holds about two jobs' buffers: jobs are held back at
    admission (and submissions refused while waiting inputs fill the
    budget are retried after a step); every job still finishes, with the
    results it has alone."""
    from raypier_optics_amd.serving import run_poisson
    jobs = _jobs(range(8), [70000] * 8)
    server = TraceServer(hbm_budget_bytes=100 << 20, max_jobs=8, max_rays_in_flight=1 << 22)
    try:
        results, _ = run_poisson(server, jobs, 1e5, seed=2)
        held, peak = server.stats()
    finally:
        server.close()
    assert held > 0 and peak > 0
    for w, r in zip(jobs[:2], results[:2]):
        a = _alone(w)
        assert r.status == 0 and (r.generations, r.interactions) == (a.generations, a.interactions)
        assert r.moments == a.moments


def test_servers_one_after_another(gpu):
    """Eight servers created, run and closed in turn in one process, each under
    a budget small enough that its buffer heap releases chunks: every server
    returns the same results, and no step stalls (with job buffers from a
    HIP memory pool per server, allocation calls from the seventh server on
    took up to seconds)."""
    import time
    jobs = _jobs(seeds=range(40, 46), sizes=[30000, 2000, 60000, 9000, 45000, 1000])
    first, worst = None, 0.0
    for k in range(8):
        # the inputs (28 MB) fit; the largest job's children (23 MB) wait for room
        server = TraceServer(hbm_budget_bytes=48 << 20, max_jobs=4, max_rays_in_flight=1 << 22)
        try:
            ids = [server.submit(w) for w in jobs]
            res = {}
            while server.busy():
                t0 = time.perf_counter()
                res.update({r.job_id: r for r in server.step()})
                worst = max(worst, time.perf_counter() - t0)
        finally:
            server.close()
        got = [res[j] for j in ids]
        if first is None:
            first = got
        for a, b in zip(first, got):
            _same(a, b, "server %d" % k)
    assert worst < 0.25, "slowest step %.3f s" % worst


def test_graph_launch_same_results(gpu):
    """A step replayed as a HIP graph (the default) and the same step
    enqueued node by node give bit-identical job results, at one and two
    generations per step."""
    jobs = _jobs(seeds=range(50, 56))
    for gens in (1, 2):
        res = []
        for graph in (False, True):
            server = TraceServer(hbm_budget_bytes=96 << 20, max_jobs=8, max_rays_in_flight=1 << 22,
                                 generations_per_step=gens, graph_launch=graph)
            try:
                ids = [server.submit(w) for w in jobs]
                done = {r.job_id: r for r in _drain(server)}
            finally:
                server.close()
            res.append([done[j] for j in ids])
        for k, (a, b) in enumerate(zip(*res)):
            _same(a, b, "job %d, %d per step" % (k, gens))
