"""GPU parity of capture selection (rpx_select_intersections behind
ctracer.select_ray_intersections / select_gausslet_intersections,
ctracer.pyx:1981-2058) against the oracle on the same device-traced inputs.

Bar: bit-exact.  The selection is a float64 face test plus copies; the
kernel and the oracle evaluate the same expressions in the same order without
contraction, so every byte of every selected record must agree.
"""
import math

import numpy as np
import pytest

import oracle
from raypier_optics_amd import _abi, workloads
from raypier_optics_amd._lib import RpxError
from raypier_optics_amd._scene import compile_scene
from raypier_optics_amd.cfaces import RectangularFace
from raypier_optics_amd.ctracer import (FaceList, Transform, RayCollection, GaussletCollection,
                                        select_ray_intersections, select_gausslet_intersections,
                                        wavelength_maps, trace_segment)
from raypier_optics_amd.tracer import trace_rays

pytestmark = pytest.mark.gpu


def tilted_plane(z, tilt_deg, half=15.0, idx=0):
    """A capture rectangle in a FaceList rotated about x and moved to z (the
    capture-plane probe's face, probes.py:97-101)."""
    face = RectangularFace(length=2 * half, width=2 * half, offset=0.0, z_plane=0.0)
    face.idx = idx
    a = math.radians(tilt_deg)
    R = np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])
    t = np.array([0.0, 0.0, z])
    fl = FaceList()
    fl.faces = [face]
    fl.transform = Transform(rotation=R, translation=t)
    fl.inverse_transform = Transform(rotation=R.T, translation=-R.T @ t)
    return fl


def oracle_select(face_set, collections, gausslet=False):
    tables = compile_scene([face_set])
    union, maps = wavelength_maps([c.wavelengths for c in collections])
    parts = [oracle.select_intersections(tables, 0, c.copy_as_array(), m, gausslet=gausslet)
             for c, m in zip(collections, maps)]
    dt = _abi.gausslet_dtype if gausslet else _abi.ray_dtype
    return (np.concatenate(parts) if parts else np.zeros(0, dt)), union


def assert_same(got, ref, what):
    assert len(got) == len(ref), (what, len(got), len(ref))
    if got.tobytes() != ref.tobytes():
        base_g = got["base_ray"] if "base_ray" in got.dtype.names else got
        base_r = ref["base_ray"] if "base_ray" in ref.dtype.names else ref
        bad = [f for f in base_r.dtype.names if base_g[f].tobytes() != base_r[f].tobytes()]
        raise AssertionError("%s: fields differ: %s" % (what, bad))


def _c2_generations(n=20000):
    w = workloads.asphere_doublet_c2(n_rays=n, seed=7)
    traced, faces = trace_rays(w.collection(), w.face_lists, recursion_limit=w.recursion_limit,
                               max_length=w.max_length)
    return w, traced


@pytest.mark.parametrize("capture", ["tilted_plane", "asphere_list"])
def test_select_rays_c2(gpu, capture):
    w, traced = _c2_generations()
    fl = tilted_plane(11.0, 10.0) if capture == "tilted_plane" else w.face_lists[0]
    before = [g.copy_as_array() for g in traced]
    out = select_ray_intersections(fl, list(traced))
    ref, union = oracle_select(fl, traced)
    assert len(ref) > 1000
    assert_same(out.copy_as_array(), ref, capture)
    assert np.array_equal(out.wavelengths, union)
    for g, b in zip(traced, before):          # the collections are not modified
        assert g.copy_as_array().tobytes() == b.tobytes()


def test_select_wavelength_union(gpu):
    """Collections with different wavelength tables: indices re-mapped into
    the sorted union (ctracer.pyx:2015-2021)."""
    w = workloads.singlet_c1(n_rays=4000)
    traced, _ = trace_rays(w.collection(), w.face_lists, recursion_limit=w.recursion_limit, max_length=w.max_length)
    arr = traced[2].copy_as_array()      # segments from the back face onwards
    rng = np.random.default_rng(3)
    ra = arr.copy()
    ra["wavelength_idx"] = rng.integers(0, 3, len(ra))
    a = RayCollection.from_array(ra)
    a.wavelengths = [0.55, 0.45, 0.65]
    rb = arr.copy()
    rb["wavelength_idx"] = rng.integers(0, 2, len(rb))
    b = RayCollection.from_array(rb)
    b.wavelengths = [0.65, 0.40]
    empty = RayCollection(0)
    empty.wavelengths = [0.70]
    fl = tilted_plane(20.0, -5.0, half=30.0)
    out = select_ray_intersections(fl, [a, empty, b])
    ref, union = oracle_select(fl, [a, empty, b])
    assert list(union) == [0.40, 0.45, 0.55, 0.65, 0.70]
    assert len(ref) > 0.3 * (len(ra) + len(rb))
    assert_same(out.copy_as_array(), ref, "wavelength union")
    got_wl = union[out.copy_as_array()["wavelength_idx"]]
    src = np.concatenate([np.array(a.wavelengths)[ra["wavelength_idx"]], np.array(b.wavelengths)[rb["wavelength_idx"]]])
    assert set(np.round(got_wl, 6)) <= set(np.round(src, 6))


def test_select_gausslets_c4(gpu):
    gc, faces, info = workloads.gausslet_c4(n_gausslets=4096, grid=16)
    traced, _ = trace_rays(gc, faces, recursion_limit=3, max_length=200.0)
    assert all(isinstance(g, GaussletCollection) for g in traced)
    fl = tilted_plane(12.0, 3.0, half=10.0)
    out = select_gausslet_intersections(fl, list(traced))
    ref, union = oracle_select(fl, traced, gausslet=True)
    assert len(ref) > 0.9 * 4096
    assert_same(out.copy_as_array(), ref, "gausslets")


def test_select_edge_cases(gpu):
    w, traced = _c2_generations(n=3000)
    gens = list(traced)
    # a face with a negative index is never selected, and cannot be traced
    # (its index would address the per-face hit counters)
    neg = tilted_plane(11.0, 10.0, idx=-1)
    assert select_ray_intersections(neg, gens).n_rays == 0
    with pytest.raises(RpxError):
        trace_segment(gens[1], [neg], [])
    # a plane no ray reaches
    far = tilted_plane(11.0, 0.0)
    far.transform = Transform(translation=(500.0, 0.0, 11.0))
    far.inverse_transform = Transform(translation=(-500.0, 0.0, -11.0))
    assert select_ray_intersections(far, gens).n_rays == 0
    # only empty collections
    e = RayCollection(0)
    e.wavelengths = [0.5]
    out = select_ray_intersections(tilted_plane(11.0, 0.0), [e, e])
    assert out.n_rays == 0 and list(out.wavelengths) == [0.5]
    # an empty list: the reference fails on ray_col_list[0]
    with pytest.raises(IndexError):
        select_ray_intersections(tilted_plane(11.0, 0.0), [])
    # a wavelength index outside its collection's table is an error, not a read
    # of another collection's table
    bad = traced[0].copy_as_array()
    bad["wavelength_idx"][5] = len(traced[0].wavelengths) + 3
    rc = RayCollection.from_array(bad)
    rc.wavelengths = traced[0].wavelengths
    with pytest.raises(RpxError):
        select_ray_intersections(tilted_plane(-2.0, 0.0), [rc])
