"""Generation consumers on the CPU: the oracle's restatements of
RayCollection._eval_neighbours (ctracer.pyx:1084-1106), the optical path
walk of MeanOpticalPathLength (results.py:112-122) and the end-face vote of
set_face_sequence (sources.py:138-155), pinned by known answers and
invariants; and the host-side vote read off an end-face histogram.
"""
import numpy as np
import pytest

import oracle
from raypier_optics_amd import _abi, workloads
from raypier_optics_amd._scene import compile_scene
from raypier_optics_amd.ctracer import RayCollection
from raypier_optics_amd.sources import _most_common, NO_FACE

REFL = _abi.REFL_RAY


def _children(parents, types):
    c = workloads.base_rays(len(parents))
    c["parent_idx"] = parents
    c["ray_type_id"] = types
    return c


def hex7():
    """A centre ray with six neighbours and a ring whose members see the
    centre and their two ring neighbours (sources.py:1063-1071 style)."""
    nb = np.full((7, 6), -1, np.int32)
    nb[0] = [1, 2, 3, 4, 5, 6]
    for k in range(6):
        r = 1 + k
        nb[r, (k + 3) % 6] = 0
        nb[r, (k + 1) % 6] = 1 + (k + 1) % 6
        nb[r, (k + 5) % 6] = 1 + (k + 5) % 6
    return nb


def test_neighbours_identity():
    """Children that are the parents' transmitted rays in parent order keep
    the parent table."""
    pnb = hex7()
    c = _children(np.arange(7), np.zeros(7))
    assert np.array_equal(oracle.eval_neighbours(c, pnb), pnb)


def test_neighbours_relabelled():
    """Each parent has one reflected and one transmitted child, shuffled: the
    child table is the parent table relabelled within each type."""
    pnb = hex7()
    rng = np.random.default_rng(2)
    parents = np.repeat(np.arange(7), 2)
    types = np.tile([0, REFL], 7)
    perm = rng.permutation(14)
    c = _children(parents[perm], types[perm] | 4)       # other type bits are ignored
    nb = oracle.eval_neighbours(c, pnb)
    for i in range(14):
        p, t = c["parent_idx"][i], c["ray_type_id"][i] & REFL
        for j in range(6):
            if pnb[p, j] < 0:
                assert nb[i, j] == -1
            else:
                k = nb[i, j]
                assert c["parent_idx"][k] == pnb[p, j] and (c["ray_type_id"][k] & REFL) == t


def test_neighbours_duplicates_and_missing():
    """Two children of one parent and type: the later one is the neighbour
    (the reference's map keeps the last assignment); a parent without a
    child of the type gives -1."""
    pnb = np.array([[1], [0], [1]], np.int32)
    c = _children([0, 1, 1, 2], [0, 0, 0, REFL])
    nb = oracle.eval_neighbours(c, pnb)
    assert nb[:, 0].tolist() == [2, 0, 0, -1]


def test_neighbours_bad_index():
    with pytest.raises(IndexError):
        oracle.eval_neighbours(_children([3], [0]), np.zeros((2, 1), np.int32))
    with pytest.raises(IndexError):
        oracle.eval_neighbours(_children([0], [0]), np.array([[5]], np.int32))


def _gen(n, parents, length, n_re, end_face=None):
    g = workloads.base_rays(n)
    g["parent_idx"] = parents
    g["length"] = length
    g["refractive_index"] = np.asarray(n_re) + 0.25j      # the imaginary part is not used
    if end_face is not None:
        g["end_face_idx"] = end_face
    return g


def test_path_lengths_known_answer():
    g0 = _gen(2, [NO_FACE] * 2, [1.0, 2.0], [1.0, 1.0])
    g1 = _gen(3, [1, 0, 1], [3.0, 4.0, 5.0], [1.5, 1.5, 1.0])
    g2 = _gen(4, [2, 0, 1, 2], [1.0, 1.0, 1.0, 2.0], [1.0, 1.0, 1.0, 1.0], [7, 7, 3, 7])
    t = oracle.path_lengths([g0, g1, g2], 7)
    assert t.tolist() == [8.0, 7.5, 9.0]
    assert len(oracle.path_lengths([g0, g1, g2], 4)) == 0


def test_path_lengths_match_accumulated_path():
    """Each interaction adds length * Re(n) of the incoming ray to the
    child's accumulated_path (cmaterials.pyx:272, 542, 923), so the walk's
    total for a ray equals its accumulated_path plus its own segment when the
    source rays start at zero."""
    w = workloads.singlet_c1(n_rays=1500)
    workloads.prepare_faces(w.face_lists, w.wavelengths, w.max_length)
    t = compile_scene(w.face_lists)
    gens, _ = oracle.trace_rays(t, w.rays, w.max_length, w.wavelengths, w.recursion_limit)
    gens = gens[:3]     # source -> front face -> back face -> detector
    det = 2     # lens faces 0, 1; detector 2
    last = gens[-1]
    tot = oracle.path_lengths(gens, det)
    sel = last[last["end_face_idx"] == det]
    assert len(tot) == len(sel) > 1000
    expect = sel["accumulated_path"] + sel["length"] * sel["refractive_index"].real
    np.testing.assert_allclose(tot, expect, rtol=1e-13)
    # dropping a term (the first generation) breaks it
    short = oracle.path_lengths(gens[1:], det)
    assert np.abs(short - expect).max() > 1.0


def test_face_sequence_vote():
    fm = {0: "a", 1: "b", 2: "c"}
    gens = [np.array([1, 1, 0, 2], np.uint32),       # 1
            np.array([2, 0, 0, 2], np.uint32),       # tie 0 / 2: the smaller id
            np.array([NO_FACE, NO_FACE, 1], np.uint32),   # no face: stop
            np.array([1], np.uint32)]
    assert oracle.face_sequence(gens, fm) == ["b", "a"]
    assert oracle.face_sequence([np.array([5, 5, 1], np.uint32)], fm) == []


@pytest.mark.parametrize("seed", range(6))
def test_histogram_vote_matches_unique(seed):
    """_most_common read off an (n_bins + 2)-bin histogram agrees with
    numpy.unique(...).argmax() (the reference's vote) on random end faces,
    including no-face and out-of-range values."""
    rng = np.random.default_rng(seed)
    n_bins = 5
    n = int(rng.integers(1, 60))
    pool = np.array([0, 1, 2, 3, 4, NO_FACE] + ([9, 17] if seed % 2 else []), np.uint32)
    ids = rng.choice(pool, size=n, p=None).astype(np.uint32)
    rays = workloads.base_rays(n)
    rays["end_face_idx"] = ids
    rc = RayCollection.from_array(rays)
    counts = np.zeros(n_bins + 2, np.uint64)
    for v in ids:
        counts[v if v < n_bins else (n_bins if v == NO_FACE else n_bins + 1)] += 1
    u, c = np.unique(ids, return_counts=True)
    assert _most_common(counts, n_bins, rc) == int(u[c.argmax()])


def test_histogram_vote_empty():
    with pytest.raises(ValueError):
        _most_common(np.zeros(4, np.uint64), 2, RayCollection(0))
