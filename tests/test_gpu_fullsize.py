"""Parity at the benchmark sizes, through properties that hold at any size.

* C2, 4,194,304 rays, float64 trace.  Rays do not interact, so the
  descendants of an input ray in the full trace are the trace of that ray
  alone: 2,000 input rays sampled at random and traced by the oracle must
  match their descendants in the full device trace (topology bit-exact,
  values within the 1e-9 parity bound).  Every generation is also
  parent-major with at most two children per parent.
* C3 with the three-order grating, 1,048,576 rays to depth 12: the same
  sampled-descendant check for 300 input rays, the ancestry followed on the
  device (the generations hold about 7e8 rays).
* C4, 262,144 gausslets x 512^2 points, float64 field sum.  The field at 24
  sampled grid points equals the oracle's sum over all gausslets (1e-9 of the
  grid's max|E|), with the same mode coefficients.
"""
import numpy as np
import pytest

import oracle
from raypier_optics_amd import workloads, cfields
from raypier_optics_amd._scene import compile_scene
from raypier_optics_amd.find_focus import find_ray_focus
from raypier_optics_amd.tracer import trace_rays
from tests.helpers import compare_generation

pytestmark = pytest.mark.gpu


def test_c2_full_size_sampled_parity(gpu):
    w = workloads.asphere_doublet_c2()
    traced, _ = trace_rays(w.collection(), w.face_lists, recursion_limit=w.recursion_limit,
                           max_length=w.max_length)
    gens = [t.copy_as_array() for t in traced]
    n0 = len(gens[0])
    assert n0 == 4194304
    for g in range(1, len(gens)):
        p = gens[g]["parent_idx"].astype(np.int64)
        assert np.all(np.diff(p) >= 0) and p[-1] < len(gens[g - 1]), "generation %d not parent-major" % g
        assert np.bincount(p).max() <= 2
    rng = np.random.default_rng(11)
    S = np.sort(rng.choice(n0, 2000, replace=False))
    # C2 rays carry their own wavelength: wavelength_idx is the input index
    assert np.array_equal(gens[0]["wavelength_idx"][S], S)
    sub = w.subset(S)
    workloads.prepare_faces(sub.face_lists, sub.wavelengths, sub.max_length)
    ref, _ = oracle.trace_rays(compile_scene(sub.face_lists), sub.rays, sub.max_length, sub.wavelengths,
                               sub.recursion_limit)
    anc = np.arange(n0)
    sel_prev = None
    for g in range(len(gens)):
        if g > 0:
            anc = anc[gens[g]["parent_idx"].astype(np.int64)]
        sel = np.nonzero(np.isin(anc, S))[0]
        if g >= len(ref):
            assert len(sel) == 0, "generation %d: sampled rays traced further on the device" % g
            continue
        got = gens[g][sel].copy()
        # renumber to the subset trace: wavelengths and parents
        got["wavelength_idx"] = np.searchsorted(S, got["wavelength_idx"]).astype(np.uint32)
        if g > 0:
            got["parent_idx"] = np.searchsorted(sel_prev, got["parent_idx"]).astype(np.uint32)
        compare_generation(got, ref[g], what="C2 full size, sampled gen %d" % g)
        sel_prev = sel
    assert len(gens) >= len(ref)


def _gather(gen, sel):
    """Host ray_dtype records of rows `sel` (device index tensor) of a
    device-resident generation."""
    from raypier_optics_amd._device import DeviceRays
    dev = gen._ensure_device()
    out = DeviceRays(max(len(sel), 1))
    out.f64()[:, :len(sel)] = dev.f64()[:, :gen.n_rays][:, sel]
    out.u32()[:, :len(sel)] = dev.u32()[:, :gen.n_rays][:, sel]
    return out.download(len(sel))


def test_c3_full_size_sampled_parity(gpu):
    """C3 with the three-order grating, 1,048,576 rays to depth 12 (about
    7e8 rays over all generations): the descendants of 300 sampled input
    rays, found on the device by following parent_idx, against the oracle's
    trace of those rays alone (topology bit-exact, values within 1e-9);
    every generation parent-major with at most three children per parent."""
    import torch
    w = workloads.michelson_c3(orders=(-1, 0, 1))
    traced, _ = trace_rays(w.collection(), w.face_lists, recursion_limit=w.recursion_limit,
                           max_length=w.max_length)
    n0 = traced[0].n_rays
    assert n0 == 1048576 and len(traced) == 12
    rng = np.random.default_rng(12)
    S = np.sort(rng.choice(n0, 300, replace=False))
    sub_rays = w.rays[S].copy()
    workloads.prepare_faces(w.face_lists, w.wavelengths, w.max_length)
    ref, _ = oracle.trace_rays(compile_scene(w.face_lists), sub_rays, w.max_length, w.wavelengths,
                               w.recursion_limit)
    S_dev = torch.as_tensor(S, device="cuda")
    anc = torch.arange(n0, device="cuda")
    sel_prev = None
    for g, gen in enumerate(traced):
        n = gen.n_rays
        if g > 0:
            parent = gen._ensure_device().u32()[1, :n].long() & 0xFFFFFFFF
            assert bool((parent[1:] >= parent[:-1]).all()), "generation %d not parent-major" % g
            assert int(torch.bincount(parent).max()) <= 3
            anc = anc[parent]
        sel = torch.nonzero(torch.isin(anc, S_dev)).flatten()
        got = _gather(gen, sel)
        assert len(got) == len(ref[g]), (g, len(got), len(ref[g]))
        assert not got["wavelength_idx"].any()                  # one wavelength: index 0 everywhere
        if g > 0:
            got["parent_idx"] = np.searchsorted(sel_prev, got["parent_idx"]).astype(np.uint32)
        compare_generation(got, ref[g], what="C3 full size, sampled gen %d" % g)
        sel_prev = sel.cpu().numpy()
    del traced
    torch.cuda.empty_cache()


def test_c4_full_size_field_samples(gpu):
    gc, faces, info = workloads.gausslet_c4()
    traced, _ = trace_rays(gc, faces, recursion_limit=3, max_length=200.0)
    out = traced[2]
    assert out.n_rays > 0.99 * 262144
    focus = find_ray_focus(out)
    pts = workloads.field_grid(focus, 512, info["span_mm"])
    modes = cfields.gausslet_modes(out)
    E = cfields.sum_gaussian_modes(out, modes, out.wavelengths, pts)
    assert E.shape == (512 * 512, 3)
    rng = np.random.default_rng(12)
    idx = rng.choice(len(pts), 24, replace=False)
    base = out.copy_as_array()["base_ray"]
    ref = oracle.sum_gaussian_modes(base, modes.cpu().numpy().view(np.complex128), np.asarray(out.wavelengths),
                                    pts[idx])
    scale = np.abs(E).max()
    assert scale > 0
    assert np.abs(E[idx] - ref).max() <= 1e-9 * scale


def test_c2_full_size_split_spot_rms(gpu):
    """The FP32-split form on the full C2 input: spot RMS per wavelength bin
    at the detector within 1e-6 of float64 and the same detector hits."""
    from raypier_optics_amd import metrics
    from raypier_optics_amd.split import SplitRayCollection, trace_rays_split
    w = workloads.asphere_doublet_c2()
    rays = w.collection()
    t64, _ = trace_rays(rays, w.face_lists, recursion_limit=w.recursion_limit, max_length=w.max_length)
    t32, _ = trace_rays_split(SplitRayCollection.from_collection(w.collection()), w.face_lists,
                              recursion_limit=w.recursion_limit, max_length=w.max_length)
    det = sum(len(fl.faces) for fl in w.face_lists) - 1
    edges = np.linspace(0.45, 0.65, 9)
    s64, n64 = metrics.spot_rms_by_wavelength(t64, det, w.wavelengths, edges)
    s32, n32 = metrics.spot_rms_by_wavelength(t32, det, w.wavelengths, edges)
    assert np.array_equal(n64, n32) and n64.sum() > 0.9 * w.n_rays
    assert np.max(np.abs(s32 - s64) / s64) <= 1e-6
