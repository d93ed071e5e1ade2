"""Compile FaceList / Face / InterfaceMaterial / Shape / Distortion objects into
the flat scene tables of include/rpx.h, and upload them to the GPU.

Objects describe themselves through `_rpx_fill_face`, `_rpx_fill_material`,
`_rpx_shape_ops` and `_rpx_compile` methods; an object without one is not
supported on the device and is refused here (RPX_E_UNSUPPORTED semantics).
"""
import collections
import ctypes
import hashlib

import numpy as np

from . import _abi


class UnsupportedOnDevice(TypeError):
    pass


class SceneTables(object):
    """Host copy of one compiled scene (numpy record arrays)."""

    def __init__(self, face_lists, faces, materials, shape_ops, vertices, zernikes, zcoefs, zinsns,
                 n_global_faces, material_objects, face_objects):
        self.face_lists = face_lists
        self.faces = faces
        self.materials = materials
        self.shape_ops = shape_ops
        self.vertices = vertices
        self.zernikes = zernikes
        self.zcoefs = zcoefs
        self.zinsns = zinsns
        self.n_global_faces = n_global_faces
        self.material_objects = material_objects
        self.face_objects = face_objects
        self._handle = None

    def desc(self):
        """rpx_scene_desc_t pointing at the host arrays (kept alive by self)."""
        def ptr(a):
            return a.ctypes.data if len(a) else None
        return _abi.SceneDesc(ptr(self.face_lists), len(self.face_lists),
                              ptr(self.faces), len(self.faces),
                              ptr(self.materials), len(self.materials),
                              ptr(self.shape_ops), len(self.shape_ops),
                              ptr(self.vertices), len(self.vertices) // 2,
                              ptr(self.zernikes), len(self.zernikes),
                              ptr(self.zcoefs), len(self.zcoefs),
                              ptr(self.zinsns), len(self.zinsns))

    @property
    def max_children(self):
        mc = {_abi.MAT_TRANSPARENT: 1, _abi.MAT_PEC: 1, _abi.MAT_DIELECTRIC: 1, _abi.MAT_GRATING: 1,
              _abi.MAT_PARTIAL: 2, _abi.MAT_FULL_DIELECTRIC: 2, _abi.MAT_FULL_DISPERSIVE: 2,
              _abi.MAT_SINGLE_LAYER: 2, _abi.MAT_COATED_DISPERSIVE: 2}
        n_list = int(sum(self.face_lists["n_faces"])) if len(self.face_lists) else 0
        best = 0
        for f in self.faces[:n_list]:
            if f["material"] >= 0:
                m = self.materials[f["material"]]
                k = mc.get(int(m["type"]), 0)
                if int(m["type"]) == _abi.MAT_GRATING:
                    k = max(1, int(m["n_orders"]))
                best = max(best, k)
        return best

    def key(self):
        """Content key: two scenes with equal tables share one upload."""
        parts = [a.tobytes() for a in (self.face_lists, self.faces, self.materials, self.shape_ops,
                                       self.vertices, self.zernikes, self.zcoefs, self.zinsns)]
        return hashlib.sha1(b"|".join(parts)).hexdigest()

    def handle(self):
        """The uploaded rpx_scene_t for these tables on the current GPU."""
        if self._handle is None:
            self._handle = _DEVICE_SCENES.get(self)
        return self._handle.ptr


class _DeviceScene(object):
    def __init__(self, tables):
        from ._lib import load, check, require_gpu
        torch = require_gpu()
        self.device = torch.cuda.current_device()
        lib = load()
        h = ctypes.c_void_p()
        d = tables.desc()
        check(lib.rpx_scene_create(ctypes.byref(d), ctypes.byref(h)))
        self.ptr = h

    def __del__(self):
        if self.ptr is not None and self.ptr.value:
            try:
                from ._lib import load
                load().rpx_scene_destroy(self.ptr)
            except Exception:
                pass
            self.ptr = None


class _SceneCache(object):
    """Recently uploaded scenes keyed by table contents and device, so that
    tracing the same optics again reuses the device copy."""

    def __init__(self, size=16):
        self.size = size
        self.entries = collections.OrderedDict()

    def get(self, tables):
        import torch
        key = (tables.key(), torch.cuda.current_device())
        ent = self.entries.get(key)
        if ent is None:
            ent = _DeviceScene(tables)
            self.entries[key] = ent
            while len(self.entries) > self.size:
                self.entries.popitem(last=False)
        else:
            self.entries.move_to_end(key)
        return ent


_DEVICE_SCENES = _SceneCache()


class SceneBuilder(object):
    def __init__(self):
        self._lists = []          # (FaceList, first_face, n_faces)
        self._faces = []          # numpy records (list faces, chain order)
        self._face_objs = []
        self._extra = []          # (record, obj) of distortion base faces
        self._materials = []
        self._mat_objs = []
        self._mat_ids = {}
        self._shape_ops = []
        self._shape_ids = {}
        self._vertices = []
        self._zernikes = []
        self._z_ids = {}
        self._zcoefs = []
        self._zinsns = []

    # ---------------------------------------------------------- helpers
    def material_index(self, mat):
        if mat is None:
            return -1
        key = id(mat)
        if key in self._mat_ids:
            return self._mat_ids[key]
        fill = getattr(mat, "_rpx_fill_material", None)
        if fill is None:
            raise UnsupportedOnDevice("material %s has no device implementation" % type(mat).__name__)
        rec = np.zeros((), _abi.material_dtype)
        fill(rec, self)
        self._materials.append(rec)
        self._mat_objs.append(mat)
        self._mat_ids[key] = len(self._materials) - 1
        return self._mat_ids[key]

    def add_vertices(self, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 2)
        if len(pts) < 1:
            raise ValueError("a polygon needs at least one vertex")
        off = len(self._vertices) // 2
        self._vertices.extend(pts.reshape(-1).tolist())
        return off, len(pts)

    def shape_root(self, shape):
        """Compile a Shape tree into postfix ops; returns the root op index."""
        if shape is None:
            return -1
        key = id(shape)
        if key in self._shape_ids:
            return self._shape_ids[key]
        begin = len(self._shape_ops)
        self._emit_shape(shape)
        root = len(self._shape_ops) - 1
        self._shape_ops[root]["begin"] = begin
        self._shape_ids[key] = root
        return root

    def _emit_shape(self, shape):
        ops = getattr(shape, "_rpx_shape_ops", None)
        if ops is None:
            raise UnsupportedOnDevice("shape %s has no device implementation" % type(shape).__name__)
        for child in getattr(shape, "_rpx_children", lambda: [])():
            self._emit_shape(child)
        rec = np.zeros((), _abi.shape_op_dtype)
        ops(rec, self)
        self._shape_ops.append(rec)

    def distortion_index(self, dist):
        key = id(dist)
        if key in self._z_ids:
            return self._z_ids[key]
        comp = getattr(dist, "_rpx_compile", None)
        if comp is None:
            raise UnsupportedOnDevice("distortion %s has no device implementation" % type(dist).__name__)
        rec = np.zeros((), _abi.zernike_dtype)
        comp(rec, self)
        self._zernikes.append(rec)
        self._z_ids[key] = len(self._zernikes) - 1
        return self._z_ids[key]

    def add_zernike_program(self, coefs, insns):
        cb = len(self._zcoefs)
        self._zcoefs.extend(coefs)
        ib = len(self._zinsns)
        self._zinsns.extend(insns)
        return cb, ib

    def extra_face(self, face):
        """A face record that belongs to no face list (distortion base)."""
        rec = self._face_record(face, -1)
        self._extra.append((rec, face))
        return len(self._extra) - 1   # patched to an absolute index in finish()

    def _face_record(self, face, list_idx):
        fill = getattr(face, "_rpx_fill_face", None)
        if fill is None:
            raise UnsupportedOnDevice("face %s has no device implementation" % type(face).__name__)
        rec = np.zeros((), _abi.face_dtype)
        rec["global_idx"] = int(face.idx)
        rec["face_list"] = list_idx
        rec["tolerance"] = float(face.tolerance)
        rec["invert_normal"] = int(face.invert_normal)
        rec["shape"] = -1
        rec["material"] = self.material_index(face.material)
        fill(rec, self)
        rec["p"][14:16] = z_band(rec, face, self)
        return rec

    # ---------------------------------------------------------- lists
    def add_face_list(self, fl):
        first = len(self._faces)
        li = len(self._lists)
        for face in fl.faces:
            self._faces.append(self._face_record(face, li))
            self._face_objs.append(face)
        self._lists.append((fl, first, len(fl.faces)))
        return li

    def finish(self):
        n_list = len(self._faces)
        faces = self._faces + [r for r, _ in self._extra]
        for r in faces:
            if r["type"] == _abi.FACE_DISTORTION and r["pad"] == 1:
                r["aux0"] = n_list + r["aux0"]
                r["pad"] = 0
        fls = np.zeros(len(self._lists), _abi.face_list_dtype)
        for i, (fl, first, n) in enumerate(self._lists):
            t = fl._trans
            it = fl._inv_trans
            fls[i]["trans"]["m"] = t[0]
            fls[i]["trans"]["t"] = t[1]
            fls[i]["inv_trans"]["m"] = it[0]
            fls[i]["inv_trans"]["t"] = it[1]
            fls[i]["first_face"] = first
            fls[i]["n_faces"] = n
        def arr(lst, dt):
            a = np.zeros(len(lst), dt)
            for i, r in enumerate(lst):
                a[i] = r
            return a
        face_arr = arr(faces, _abi.face_dtype)
        n_glob = int(face_arr["global_idx"][:n_list].max()) + 1 if n_list else 0
        return SceneTables(fls, face_arr,
                           arr(self._materials, _abi.material_dtype),
                           arr(self._shape_ops, _abi.shape_op_dtype),
                           np.asarray(self._vertices, dtype=np.float64),
                           arr(self._zernikes, _abi.zernike_dtype),
                           arr(self._zcoefs, _abi.zernike_coef_dtype),
                           arr(self._zinsns, _abi.zernike_insn_dtype),
                           n_glob, list(self._mat_objs), list(self._face_objs))


def z_band(rec, face, builder=None):
    """(lo, hi): a z range, in the face list's frame, holding every point at
    which the face can accept a base-ray hit, widened by a margin for
    rounding (the root of a near-tangent sphere intersection can sit about
    1e-6 mm off the surface); (nan, nan) where this build does not bound it.
    The kernels skip a face whose band lies outside the part of a ray's
    segment closer than its nearest hit so far.

    A DistortionFace over a planar base face gets the plane's z widened by a
    bound on the distortion's sag over the face's aperture (the sum over its
    Zernike terms of |N c| times the radial polynomial's absolute
    coefficients at the aperture radius) and by 1e-3 mm plus twice that
    bound for its fixed-point iteration, whose accepted point lies on or
    next to the distorted surface (about 3.5e-4 mm from the plane on the
    C3 mirror)."""
    import math
    from .cshapes import bound_radius
    nan = (float("nan"), float("nan"))
    t = int(rec["type"])
    p = rec["p"].tolist()
    if t in (_abi.FACE_CIRCULAR, _abi.FACE_SHAPED_PLANAR, _abi.FACE_RECTANGULAR, _abi.FACE_POLYGON,
             _abi.FACE_EXTRUDED_PLANAR):
        if t == _abi.FACE_EXTRUDED_PLANAR:
            lo, hi = p[4], p[5]
        else:
            z = {_abi.FACE_CIRCULAR: p[2], _abi.FACE_RECTANGULAR: p[3]}.get(t, p[0])
            lo = hi = z
        m = 1e-7 * (1.0 + abs(lo) + abs(hi))
    elif t in (_abi.FACE_SPHERICAL, _abi.FACE_SHAPED_SPHERICAL):
        if t == _abi.FACE_SPHERICAL:
            R, z, r = p[1], p[2], abs(p[0]) / 2
        else:
            R, z, r = p[0], p[1], bound_radius(face.shape)
        if R == 0.0 or not math.isfinite(R) or math.isnan(r):
            return nan
        aR = abs(R)
        r = r * (1.0 + 1e-6)
        depth = aR - math.sqrt(aR * aR - r * r) if r < aR else aR
        lo, hi = (z - depth, z) if R > 0 else (z, z + depth)
        m = 1e-6 * (1.0 + abs(z) + aR + min(r, aR))
    elif t == _abi.FACE_ASPHERIC:
        band = p[12]
        if band < 0:
            return nan
        lo, hi = p[1] - band, p[1] + band
        m = 1e-9 * (1.0 + abs(p[1]) + band)
    elif t == _abi.FACE_DISTORTION and builder is not None:
        base_rec = builder._extra[int(rec["aux0"])][0]
        bt = int(base_rec["type"])
        if bt not in (_abi.FACE_CIRCULAR, _abi.FACE_SHAPED_PLANAR, _abi.FACE_RECTANGULAR, _abi.FACE_POLYGON):
            return nan
        bp = base_rec["p"].tolist()
        z = {_abi.FACE_CIRCULAR: bp[2], _abi.FACE_RECTANGULAR: bp[3]}.get(bt, bp[0])
        zrec = builder._zernikes[int(rec["aux1"])]
        rho = bound_radius(face.shape) / float(zrec["unit_radius"])
        if not math.isfinite(rho):
            return nan
        rho = max(rho, 1.0)
        sag = 0.0
        for c in builder._zcoefs[int(zrec["coef_begin"]): int(zrec["coef_begin"]) + int(zrec["n_coefs"])]:
            deg = int(c["deg"][0])
            sag += abs(float(c["nval"])) * sum(abs(float(a)) * rho ** i for i, a in enumerate(c["poly"][0][:deg + 1]))
        if not math.isfinite(sag):
            return nan
        lo, hi = z - sag, z + sag
        m = 1e-3 + 2.0 * sag + 1e-7 * (1.0 + abs(z))
    else:
        return nan
    if not (math.isfinite(lo) and math.isfinite(hi)):
        return nan
    return lo - m, hi + m


def compile_scene(face_sets):
    b = SceneBuilder()
    for fl in face_sets:
        b.add_face_list(fl)
    return b.finish()
