"""CPU oracle for the raypier tracing loop and gausslet field sum.

TEST INFRASTRUCTURE ONLY.  This package (rpx_oracle.c restating the
reference Cython statement by statement, built by oracle/Makefile) is imported
by tests/, by __graft_entry__.smoke() as the checker and by bench.py's
cpu_baseline leg.  The product library never links or calls it.

Functions take the scene tables compiled by raypier_optics_amd._scene and
numpy arrays of the packed ray_dtype / gausslet_dtype records.

Parity unpinned (checked only against this restatement): the libm
boundary, the DistortionFace result for rays that have not converged in its
20 fixed-point steps, and the Zernike derivative values for n >= 5; in
oracle/sources.py and oracle/fields.py, `exp` in the source amplitude weights.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        vp, u64, i32, f64 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32, ctypes.c_double
        lib.oracle_trace_segment.restype = u64
        lib.oracle_trace_segment.argtypes = [vp, i32, i32, vp, u64, ctypes.c_float, vp, u64, vp, u64, vp]
        lib.oracle_trace_segment_mt.restype = u64
        lib.oracle_trace_segment_mt.argtypes = [vp, i32, i32, vp, u64, ctypes.c_float, vp, u64, vp, u64, vp, u64,
                                                i32, i32]
        lib.oracle_trace_gausslet.restype = u64
        lib.oracle_trace_gausslet.argtypes = [vp, i32, i32, vp, u64, f64, vp, u64, vp, u64, vp]
        lib.oracle_face_intersect.restype = f64
        lib.oracle_face_intersect.argtypes = [vp, i32, vp, vp, i32]
        lib.oracle_face_normal.argtypes = [vp, i32, vp, vp]
        lib.oracle_shape_inside.restype = i32
        lib.oracle_shape_inside.argtypes = [vp, i32, f64, f64]
        lib.oracle_zernike.argtypes = [vp, i32, vp, vp, u64, vp, vp]
        lib.oracle_eval_child_ray.restype = u64
        lib.oracle_eval_child_ray.argtypes = [vp, i32, vp, ctypes.c_uint32, vp, vp, vp, vp, u64, vp, u64]
        lib.oracle_dispersion.argtypes = [vp, vp, u64, vp]
        lib.oracle_set_f32_geometry.restype = None
        lib.oracle_set_f32_geometry.argtypes = [i32]
        lib.oracle_evaluate_modes.argtypes = [vp, vp, vp, vp, u64, i32, f64, vp]
        lib.oracle_neighbours_gc.argtypes = [vp, u64, vp, vp, vp, vp]
        lib.oracle_sum_gaussian_modes.argtypes = [vp, u64, vp, vp, u64, vp, u64, f64, vp]
        lib.oracle_select_intersections.restype = u64
        lib.oracle_select_intersections.argtypes = [vp, i32, vp, u64, i32, vp, vp]
        lib.oracle_eval_neighbours.restype = i32
        lib.oracle_eval_neighbours.argtypes = [vp, vp, u64, vp, u64, i32, vp, vp]
        lib.oracle_config_parabasal_rays.restype = i32
        lib.oracle_config_parabasal_rays.argtypes = [vp, u64, vp, u64, f64, f64]
        lib.oracle_project_to_plane.argtypes = [vp, u64, vp, vp]
        lib.oracle_lagrange_invariant.restype = i32
        lib.oracle_lagrange_invariant.argtypes = [vp, u64, vp, u64, vp]
        lib.oracle_calc_mode_curvature.argtypes = [vp] * 5 + [u64] + [vp] * 6
        lib.oracle_apply_mode_curvature.argtypes = [vp, u64, vp, vp, vp]
        lib.oracle_build_interaction_matrix.restype = u64
        lib.oracle_build_interaction_matrix.argtypes = [vp, vp, u64, vp, vp, u64] + [vp] * 6 + [f64] * 4 + [vp] * 3
        lib.oracle_sizes.restype = i32
        lib.oracle_sizes.argtypes = [i32]
        _lib = lib
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None and a.size else None


class _Desc(object):
    """Keeps a scene's host tables alive while the oracle reads them."""

    def __init__(self, tables):
        self.tables = tables
        self.desc = tables.desc()

    @property
    def ref(self):
        return ctypes.byref(self.desc)


def trace_segment(tables, rays, max_length, wavelengths, one_list=-1, one_face=-1, max_children=3, threads=1):
    """One generation; `rays` (ray_dtype) is updated in place.  Returns
    (children, face_hits).  threads > 1 runs the OpenMP form (same
    output)."""
    lib = load()
    d = _Desc(tables)
    wl = np.ascontiguousarray(wavelengths, dtype=np.float64)
    cap = max_children * len(rays) + 1
    out = np.empty(cap, rays.dtype) if threads > 1 else np.zeros(cap, rays.dtype)
    hits = np.zeros(max(tables.n_global_faces, 1), np.uint32)
    if threads > 1:
        n = lib.oracle_trace_segment_mt(d.ref, one_list, one_face, _p(rays), len(rays), max_length, _p(wl), len(wl),
                                        _p(out), cap, _p(hits), len(hits), max_children, threads)
    else:
        n = lib.oracle_trace_segment(d.ref, one_list, one_face, _p(rays), len(rays), max_length, _p(wl), len(wl),
                                     _p(out), cap, _p(hits))
    assert n <= cap
    return out[:n].copy(), hits


def select_intersections(tables, face_list, rays, wl_map, gausslet=False):
    """Copies of the rays (gausslets) of one collection whose segment meets
    face list `face_list`, wavelength_idx mapped through `wl_map`."""
    lib = load()
    d = _Desc(tables)
    rays = np.ascontiguousarray(rays)
    m = np.ascontiguousarray(wl_map, dtype=np.uint32)
    out = np.zeros(len(rays) + 1, rays.dtype)
    n = lib.oracle_select_intersections(d.ref, face_list, _p(rays), len(rays), int(bool(gausslet)), _p(m), _p(out))
    return out[:n].copy()


def eval_neighbours(children, parent_nb):
    """Neighbour table of a child generation (ray_dtype array) from its
    parent generation's (n_parent, K) table."""
    lib = load()
    pnb = np.ascontiguousarray(parent_nb, dtype=np.int32)
    n_par, k = pnb.shape
    pidx = np.ascontiguousarray(children["parent_idx"], dtype=np.uint32)
    rtype = np.ascontiguousarray(children["ray_type_id"], dtype=np.uint32)
    nb = np.zeros((len(children), k), np.int32)
    rmap = np.zeros(max(2 * n_par, 1), np.int32)
    if lib.oracle_eval_neighbours(_p(pidx), _p(rtype), len(children), _p(pnb), n_par, k, _p(rmap), _p(nb)):
        raise IndexError("parent or neighbour index outside the parent generation")
    return nb


def path_lengths(generations, idx):
    """MeanOpticalPathLength._calc_result (results.py:112-122) before the
    mean: `generations` are ray_dtype arrays in trace order."""
    all_rays = list(reversed(generations))
    last = all_rays[0]
    selected_idx = np.argwhere(last["end_face_idx"] == idx).ravel()
    total = np.zeros(len(selected_idx), "d")
    for ray in all_rays:
        selected = ray[selected_idx]
        total += selected["length"] * selected["refractive_index"].real
        selected_idx = selected["parent_idx"]
    return total


def face_sequence(end_face_idx_per_generation, face_map):
    """set_face_sequence (sources.py:143-155) over the end_face_idx arrays of
    the generations; face_map maps a global face index to its entry."""
    out = []
    for face_ids in end_face_idx_per_generation:
        ids, counts = np.unique(face_ids, return_counts=True)
        most_common = ids[counts.argmax()]
        try:
            out.append(face_map[most_common])
        except KeyError:
            break
    return out


def config_parabasal_rays(gausslets, wavelength_list, radius, working_dist):
    """A copy of the gausslet_dtype array with its parabasal rays configured."""
    lib = load()
    g = np.ascontiguousarray(gausslets).copy()
    wl = np.ascontiguousarray(wavelength_list, dtype=np.float64)
    if lib.oracle_config_parabasal_rays(_p(g), len(g), _p(wl), len(wl), float(radius), float(working_dist)):
        raise IndexError("wavelength_idx outside the wavelength list")
    return g


def project_to_plane(gausslets, origin, direction):
    lib = load()
    g = np.ascontiguousarray(gausslets).copy()
    o = np.ascontiguousarray(origin, dtype=np.float64)
    d = np.ascontiguousarray(direction, dtype=np.float64)
    lib.oracle_project_to_plane(_p(g), len(g), _p(o), _p(d))
    return g


def lagrange_invariant(gausslets, wavelengths):
    """GaussletCollection.lagrange_invariant of a gausslet_dtype array."""
    lib = load()
    g = np.ascontiguousarray(gausslets)
    wl = np.ascontiguousarray(wavelengths, dtype=np.float64)
    out = np.zeros(len(g))
    if lib.oracle_lagrange_invariant(_p(g), len(g), _p(wl), len(wl), _p(out)):
        raise IndexError("wavelength_idx outside the wavelength list")
    return out


def calc_mode_curvature(rx, ry, dx, dy, dx2, dy2, dxdy):
    lib = load()
    ins = [np.ascontiguousarray(v, np.float64) for v in (dx, dy, dx2, dy2, dxdy)]
    n = len(ins[0])
    A, B, C = (np.zeros(n) for _ in range(3))
    X, Y, Z = (np.zeros((n, 3)) for _ in range(3))
    lib.oracle_calc_mode_curvature(*[_p(v) for v in ins], n, _p(A), _p(B), _p(C), _p(X), _p(Y), _p(Z))
    return A, B, C, X, Y, Z


def apply_mode_curvature(gausslets, A, B, C):
    """A copy of the gausslet_dtype array with the curvature applied."""
    lib = load()
    g = np.ascontiguousarray(gausslets).copy()
    A, B, C = (np.ascontiguousarray(v, np.float64) for v in (A, B, C))
    lib.oracle_apply_mode_curvature(_p(g), len(g), _p(A), _p(B), _p(C))
    return g


def build_interaction_matrix(rx, ry, tx, ty, A, B, C, x, y, z, wavelength, spacing, max_spacing, blending):
    lib = load()
    rx, ry, tx, ty, A, B, C = (np.ascontiguousarray(v, np.float64) for v in (rx, ry, tx, ty, A, B, C))
    x, y, z = (np.ascontiguousarray(v, np.float64).reshape(-1, 3) for v in (x, y, z))
    n, l = len(rx), len(tx)
    data = np.zeros(max(n * l, 1), np.complex128)
    xi = np.zeros(max(n * l, 1), np.uint64)
    xj = np.zeros(max(n * l, 1), np.uint64)
    m = lib.oracle_build_interaction_matrix(_p(rx), _p(ry), n, _p(tx), _p(ty), l, _p(A), _p(B), _p(C), _p(x), _p(y),
                                            _p(z), wavelength, spacing, max_spacing, blending, _p(data), _p(xi),
                                            _p(xj))
    return data[:m].copy(), (xi[:m].copy(), xj[:m].copy())


def trace_gausslet(tables, gs, max_length, wavelengths, one_list=-1, one_face=-1, max_children=3):
    lib = load()
    d = _Desc(tables)
    wl = np.ascontiguousarray(wavelengths, dtype=np.float64)
    cap = max_children * len(gs) + 1
    out = np.zeros(cap, gs.dtype)
    hits = np.zeros(max(tables.n_global_faces, 1), np.uint32)
    n = lib.oracle_trace_gausslet(d.ref, one_list, one_face, _p(gs), len(gs), max_length, _p(wl), len(wl),
                                  _p(out), cap, _p(hits))
    assert n <= cap
    return out[:n].copy(), hits


def trace_rays(tables, rays, max_length, wavelengths, recursion_limit, gausslet=False, threads=1):
    """Generation loop of core/tracer.py:9-47 over an already compiled scene
    (ray generations on `threads` OpenMP threads when threads > 1)."""
    rays = rays.copy()
    if gausslet:
        rays["base_ray"]["length"] = max_length
        rays["para_rays"]["length"] = max_length
    else:
        rays["length"] = max_length
    traced = []
    count = 0
    hits = np.zeros(max(tables.n_global_faces, 1), np.uint64)
    while len(rays) > 0 and count < recursion_limit:
        traced.append(rays)
        if gausslet:
            rays, h = trace_gausslet(tables, rays, max_length, wavelengths)
        else:
            rays, h = trace_segment(tables, rays, max_length, wavelengths, threads=threads)
        hits += h
        count += 1
    return traced, hits


def face_intersect(tables, rec, p1, p2, is_base_ray):
    lib = load()
    d = _Desc(tables)
    a = np.ascontiguousarray(p1, np.float64)
    b = np.ascontiguousarray(p2, np.float64)
    return lib.oracle_face_intersect(d.ref, rec, _p(a), _p(b), int(is_base_ray))


def face_normal(tables, rec, p):
    lib = load()
    d = _Desc(tables)
    a = np.ascontiguousarray(p, np.float64)
    out = np.zeros(3)
    lib.oracle_face_normal(d.ref, rec, _p(a), _p(out))
    return out


def shape_inside(tables, root, x, y):
    lib = load()
    d = _Desc(tables)
    return bool(lib.oracle_shape_inside(d.ref, root, float(x), float(y)))


def zernike(tables, zi, x, y):
    lib = load()
    d = _Desc(tables)
    x = np.ascontiguousarray(x, np.float64).ravel()
    y = np.ascontiguousarray(y, np.float64).ravel()
    sag = np.zeros(len(x))
    grad = np.zeros((len(x), 3))
    lib.oracle_zernike(d.ref, zi, _p(x), _p(y), len(x), _p(sag), _p(grad))
    return sag, grad


def eval_child_ray(tables, mat, ray, idx, point, normal, tangent, wavelengths=()):
    """Children of one ray (ray_dtype record array of length 1, mutated as the
    reference mutates nothing here)."""
    lib = load()
    d = _Desc(tables)
    r = np.ascontiguousarray(ray).copy()
    wl = np.ascontiguousarray(wavelengths, np.float64)
    out = np.zeros(4, r.dtype)
    pt, nm, tg = (np.ascontiguousarray(v, np.float64) for v in (point, normal, tangent))
    n = lib.oracle_eval_child_ray(d.ref, mat, _p(r), idx, _p(pt), _p(nm), _p(tg), _p(wl) if len(wl) else None,
                                  len(wl), _p(out), 4)
    return out[:n].copy()


def dispersion(disp_rec, wl):
    lib = load()
    rec = np.ascontiguousarray(disp_rec)
    w = np.ascontiguousarray(wl, np.float64)
    out = np.zeros(2 * len(w))
    lib.oracle_dispersion(_p(rec), _p(w), len(w), _p(out))
    return out[0::2] + 1j * out[1::2]


def evaluate_modes(x, y, dx, dy, blending=1.0):
    lib = load()
    x, y, dx, dy = (np.ascontiguousarray(v, np.float64) for v in (x, y, dx, dy))
    n, size = x.shape
    out = np.zeros((n, 3), np.complex128)
    lib.oracle_evaluate_modes(_p(x), _p(y), _p(dx), _p(dy), n, size, blending, _p(out))
    return out


def neighbours_gc(gs):
    lib = load()
    g = np.ascontiguousarray(gs)
    n = len(g)
    X, Y, DX, DY = (np.zeros((n, 6)) for _ in range(4))
    lib.oracle_neighbours_gc(_p(g), n, _p(X), _p(Y), _p(DX), _p(DY))
    return X, Y, DX, DY


def sum_gaussian_modes(rays, modes, wavelengths, points, time_ps=0.0):
    lib = load()
    r = np.ascontiguousarray(rays)
    m = np.ascontiguousarray(modes, np.complex128)
    wl = np.ascontiguousarray(wavelengths, np.float64)
    pts = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    out = np.zeros((len(pts), 3), np.complex128)
    lib.oracle_sum_gaussian_modes(_p(r), len(r), _p(m), _p(wl), len(wl), _p(pts), len(pts), time_ps, _p(out))
    return out


def _round_f32(a):
    """Copy of a record array with every finite float64 field (nested ones
    too) within float32's range rounded to float32."""
    a = a.copy()

    def walk(arr):
        for name in arr.dtype.names or ():
            sub = arr[name]
            if sub.dtype.names:
                walk(sub)
            elif sub.dtype.base == np.float64:
                v = np.asarray(sub)
                ok = np.isfinite(v) & (np.abs(v) < 3e38)
                v2 = np.where(ok, v.astype(np.float32).astype(np.float64), v)
                arr[name] = v2
    walk(a)
    return a


def trace_rays_f32_geometry(tables, rays, max_length, wavelengths, recursion_limit, threads=1):
    """trace_rays with the oracle's float32-geometry emulation: a
    measurement of the precision form that holds ray geometry, face
    parameters and refractive indices in float32 and keeps float64 for the
    root solves (run here in float64 on the float32-rounded operands),
    optical path and phase.  Face lists, faces and the materials' fixed
    indices are rounded to float32; so are the input rays' fields other than
    phase and optical path.  The emulation rounds each result once where the
    form stores or computes it in float32, so it understates that form's
    per-ray rounding (several float32 operations per result), never its
    systematic part (float32 parameters and indices)."""
    import copy
    t = copy.copy(tables)
    t.face_lists = _round_f32(tables.face_lists)
    t.faces = _round_f32(tables.faces)
    mats = tables.materials.copy()
    for name in ("n_inside", "n_outside"):
        if name in (mats.dtype.names or ()):
            mats[name] = mats[name].astype(np.float32).astype(np.float64)
    t.materials = mats
    r = rays.copy()
    keep = {"phase", "accumulated_path"}
    for name in r.dtype.names:
        base = r.dtype[name].base
        if name in keep:
            continue
        if base == np.float64:
            v = r[name]
            r[name] = np.where(np.isfinite(v), v.astype(np.float32).astype(np.float64), v)
        elif base == np.complex128:
            v = r[name]
            r[name] = v.real.astype(np.float32).astype(np.float64) + 1j * v.imag.astype(np.float32).astype(np.float64)
    lib = load()
    lib.oracle_set_f32_geometry(1)
    try:
        return trace_rays(t, r, max_length, wavelengths, recursion_limit, threads=threads)
    finally:
        lib.oracle_set_f32_geometry(0)
