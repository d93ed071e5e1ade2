"""Capture selection (select_ray_intersections / select_gausslet_intersections,
ctracer.pyx:1981-2058) on the CPU: the oracle's restatement pinned against
closed forms on a capture rectangle, and the host-side wavelength re-indexing
against the defining property of np.unique's inverse.

Closed forms: a rectangle (RectangularFace, cfaces.pyx:354-397: |x - offset|
<= length/2, |y| <= width/2) in the plane z = 10 of a translated FaceList;
a ray from z = 0 with direction d meets it at distance 10 / d_z when its
segment is at least that long.
"""
import math

import numpy as np
import pytest

import oracle
from raypier_optics_amd import _abi, workloads
from raypier_optics_amd._scene import compile_scene
from raypier_optics_amd.cfaces import RectangularFace
from raypier_optics_amd.ctracer import FaceList, Transform, wavelength_maps


def capture_list(z=10.0, idx=0):
    face = RectangularFace(length=4.0, width=6.0, offset=0.0, z_plane=0.0)
    face.idx = idx
    fl = FaceList()
    fl.faces = [face]
    fl.transform = Transform(translation=(0.0, 0.0, z))
    fl.inverse_transform = Transform(translation=(0.0, 0.0, -z))
    return fl


def rays_toward_plane():
    """(rays, expected selected indices, expected lengths)."""
    r = workloads.base_rays(7)
    d_slant = np.array([0.1, 0.0, 1.0]) / math.sqrt(1.01)
    spec = [
        ((1.0, 1.0, 0.0), (0, 0, 1), 100.0),     # 0: hit at 10
        ((1.0, 1.0, 0.0), (0, 0, 1), 5.0),       # 1: segment ends before the plane
        ((2.5, 0.0, 0.0), (0, 0, 1), 100.0),     # 2: |x| > length/2: miss
        ((0.0, 2.5, 0.0), (0, 0, 1), 100.0),     # 3: |y| <= width/2: hit
        ((0.0, 0.0, 0.0), tuple(d_slant), 100.0),  # 4: slanted hit at 10*sqrt(1.01)
        ((0.0, 0.0, 12.0), (0, 0, 1), 100.0),    # 5: starts past the plane
        ((0.0, 0.0, 20.0), (0, 0, -1), 100.0),   # 6: from above: hit at 10
    ]
    for i, (o, d, L) in enumerate(spec):
        r["origin"][i] = o
        r["direction"][i] = d
        r["length"][i] = L
        r["wavelength_idx"][i] = i % 2
        r["end_face_idx"][i] = 77
    sel = [0, 3, 4, 6]
    lengths = [10.0, 10.0, 10.0 * math.sqrt(1.01), 10.0]
    return r, sel, lengths


def test_oracle_select_closed_form():
    rays, sel, lengths = rays_toward_plane()
    before = rays.copy()
    t = compile_scene([capture_list()])
    wl_map = np.array([5, 3], np.uint32)
    out = oracle.select_intersections(t, 0, rays, wl_map)
    assert rays.tobytes() == before.tobytes()        # the input is not modified
    assert len(out) == len(sel)
    np.testing.assert_allclose(out["length"], lengths, rtol=1e-13)
    assert np.array_equal(out["end_face_idx"], np.zeros(len(sel), np.uint32))
    assert np.array_equal(out["wavelength_idx"], wl_map[before["wavelength_idx"][sel]])
    for f in ("origin", "direction", "E1_amp", "phase", "accumulated_path", "ray_type_id", "parent_idx"):
        assert np.array_equal(out[f], before[f][sel]), f


def test_oracle_select_face_index():
    """The copy's end_face_idx is the face's global index; a face whose index
    is negative is hit (the length still shrinks on the copy) but the ray is
    not selected (ctracer.pyx:2008-2011)."""
    rays, sel, _ = rays_toward_plane()
    t = compile_scene([capture_list(idx=9)])
    out = oracle.select_intersections(t, 0, rays, np.arange(2, dtype=np.uint32))
    assert np.array_equal(out["end_face_idx"], np.full(len(sel), 9, np.uint32))
    t = compile_scene([capture_list(idx=-1)])
    assert len(oracle.select_intersections(t, 0, rays, np.arange(2, dtype=np.uint32))) == 0


def test_oracle_select_gausslets():
    """Gausslets: the base ray decides and is updated; the parabasal rays are
    copied unchanged (ctracer.pyx:2042-2051)."""
    rays, sel, lengths = rays_toward_plane()
    g = np.zeros(len(rays), _abi.gausslet_dtype)
    g["base_ray"] = rays
    rng = np.random.default_rng(4)
    for f in ("origin", "direction", "normal"):
        g["para_rays"][f] = rng.normal(size=g["para_rays"][f].shape)
    g["para_rays"]["length"] = rng.uniform(1, 2, g["para_rays"]["length"].shape)
    t = compile_scene([capture_list()])
    out = oracle.select_intersections(t, 0, g, np.array([1, 0], np.uint32), gausslet=True)
    assert len(out) == len(sel)
    assert out["para_rays"].tobytes() == g["para_rays"][sel].tobytes()
    np.testing.assert_allclose(out["base_ray"]["length"], lengths, rtol=1e-13)
    assert np.array_equal(out["base_ray"]["wavelength_idx"], 1 - rays["wavelength_idx"][sel])


def test_oracle_select_empty():
    t = compile_scene([capture_list()])
    assert len(oracle.select_intersections(t, 0, workloads.base_rays(0), np.zeros(1, np.uint32))) == 0


@pytest.mark.parametrize("tables", [
    [[0.6, 0.5], [0.5, 0.7]],
    [[1.064], [], [0.4, 1.064, 0.3]],
    [[0.5, 0.5, 0.2]],
])
def test_wavelength_maps(tables):
    """union is sorted and duplicate-free, and union[map_t[j]] is wavelength j
    of table t -- what np.unique's inverse means for the concatenation."""
    union, maps = wavelength_maps(tables)
    flat = np.concatenate([np.asarray(w, np.float64) for w in tables])
    assert np.array_equal(union, np.unique(flat))
    assert np.all(np.diff(union) > 0)
    assert [len(m) for m in maps] == [len(w) for w in tables]
    for w, m in zip(tables, maps):
        assert m.dtype == np.uint32
        assert np.array_equal(union[m], np.asarray(w, np.float64))


def test_wavelength_maps_known():
    union, maps = wavelength_maps([[0.6, 0.5], [0.5, 0.7]])
    assert list(union) == [0.5, 0.6, 0.7]
    assert [list(m) for m in maps] == [[1, 0], [0, 2]]
