"""Traces from several host threads at once, each on a stream of its own.

The library keeps per-thread state (last error string, timing events, the
pinned block the generation control head is read back into), and the
single-pass generation kernels of different streams share the device.  Every
thread's generations must be bit-identical to a serial trace of the same
input.
"""
import threading

import pytest

from raypier_optics_amd import workloads
from raypier_optics_amd.tracer import trace_rays
from raypier_optics_amd.split import trace_rays_split

pytestmark = pytest.mark.gpu

SEEDS = (21, 22, 23, 24)


def _trace(seed, split):
    w = workloads.asphere_doublet_c2(n_rays=65536, seed=seed)
    tracer = trace_rays_split if split else trace_rays
    traced, _ = tracer(w.collection(), w.face_lists, recursion_limit=w.recursion_limit, max_length=w.max_length)
    if split:
        traced = [t.to_collection() for t in traced]
    return [t.copy_as_array() for t in traced]


@pytest.mark.parametrize("split", [False, True])
def test_threads_match_serial(gpu, split):
    import torch
    ref = {s: _trace(s, split) for s in SEEDS}
    got, errors = {}, []

    def run(seed):
        try:
            stream = torch.cuda.Stream()
            with torch.cuda.stream(stream):
                got[seed] = _trace(seed, split)
            stream.synchronize()
        except Exception as e:      # reported below, in the main thread
            errors.append("seed %d: %r" % (seed, e))

    threads = [threading.Thread(target=run, args=(s,)) for s in SEEDS]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in threads), "a tracing thread did not finish"
    assert not errors, errors
    for s in SEEDS:
        assert len(got[s]) == len(ref[s]), "seed %d: %d vs %d generations" % (s, len(got[s]), len(ref[s]))
        for g, (a, b) in enumerate(zip(got[s], ref[s])):
            assert a.tobytes() == b.tobytes(), "seed %d generation %d differs from the serial trace" % (s, g)
