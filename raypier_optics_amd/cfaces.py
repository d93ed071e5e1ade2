"""Faces (drop-in for the device-supported subset of raypier.core.cfaces).

Each class keeps the reference's constructor keywords and attributes and
compiles to one rpx_face_t record; intersection and normals run on the GPU
(rpx_faces.h).  Parameter slots follow include/rpx.h.
"""
import math

import numpy as np

from . import _abi
from .ctracer import Face, Shape, norm


class ShapedFace(Face):
    """Face with an aperture Shape (cfaces.pyx:50-137)."""

    def __init__(self, **kwds):
        super().__init__(**{k: v for k, v in kwds.items() if k in ("owner", "tolerance", "max_length",
                                                                     "material", "invert_normal")})
        self.shape = kwds.get("shape", Shape())
        self.invert_normals = int(kwds.get("invert_normals", 0))

    def _fill_shape(self, rec, builder):
        rec["shape"] = builder.shape_root(self.shape)
        rec["invert_normals"] = int(self.invert_normals)


def _base_kwds(kwds):
    return {k: v for k, v in kwds.items() if k in ("owner", "tolerance", "max_length", "material",
                                                    "invert_normal")}


class CircularFace(Face):
    """Circular plane z = z_plane (cfaces.pyx:140-191)."""
    params = ['diameter', 'offset']

    def __init__(self, **kwds):
        super().__init__(**_base_kwds(kwds))
        self.diameter = kwds.get("diameter", 0.0)
        self.offset = kwds.get("offset", 0.0)
        self.z_plane = kwds.get("z_plane", 0.0)
        self.invert_normals = kwds.get("invert_normals", False)

    def _rpx_fill_face(self, rec, b):
        rec["type"] = _abi.FACE_CIRCULAR
        rec["p"][:3] = (self.diameter, self.offset, self.z_plane)
        rec["invert_normals"] = int(bool(self.invert_normals))


class ShapedPlanarFace(ShapedFace):
    """Plane z = z_height bounded by a Shape (cfaces.pyx:194-242)."""

    def __init__(self, **kwds):
        super().__init__(**kwds)
        self.z_height = kwds.get("z_height", 0.0)

    def _rpx_fill_face(self, rec, b):
        rec["type"] = _abi.FACE_SHAPED_PLANAR
        rec["p"][0] = self.z_height
        self._fill_shape(rec, b)

    def eval_z(self, x, y):
        return self.z_height


class RectangularFace(Face):
    """Rectangle in z = z_plane (cfaces.pyx:354-407)."""
    params = ['length', 'width', 'offset']

    def __init__(self, **kwds):
        super().__init__(**_base_kwds(kwds))
        self.z_plane = kwds.get("z_plane", 0.0)
        self.width = kwds.get("width", 2.0)
        self.length = kwds.get("length", 5.0)
        self.offset = kwds.get("offset", 0.0)

    def _rpx_fill_face(self, rec, b):
        rec["type"] = _abi.FACE_RECTANGULAR
        rec["p"][:4] = (self.length, self.width, self.offset, self.z_plane)


class PolygonFace(Face):
    """Polygon in z = z_plane (cfaces.pyx:1077-1118)."""

    def __init__(self, z_plane=0.0, xy_points=((),), **kwds):
        super().__init__(**_base_kwds(kwds))
        self.z_plane = z_plane
        self.xy_points = xy_points

    @property
    def xy_points(self):
        return self._xy_points

    @xy_points.setter
    def xy_points(self, pts):
        self._xy_points = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 2)

    def _rpx_fill_face(self, rec, b):
        rec["type"] = _abi.FACE_POLYGON
        rec["p"][0] = self.z_plane
        off, n = b.add_vertices(self._xy_points)
        rec["aux0"] = off
        rec["aux1"] = n


class SphericalFace(Face):
    """Spherical cap with diameter aperture (cfaces.pyx:410-498).

    curvature = z_height - z_centre."""
    params = ['diameter']

    def __init__(self, **kwds):
        super().__init__(**_base_kwds(kwds))
        self.diameter = kwds.get("diameter", 0.0)
        self.z_height = kwds.get("z_height", 0.0)
        self.curvature = kwds.get("curvature", 25.0)

    def _rpx_fill_face(self, rec, b):
        rec["type"] = _abi.FACE_SPHERICAL
        rec["p"][:3] = (self.diameter, self.curvature, self.z_height)


class ShapedSphericalFace(ShapedFace):
    """Spherical cap bounded by a Shape (cfaces.pyx:501-608)."""
    params = []

    def __init__(self, **kwds):
        super().__init__(**kwds)
        self.z_height = kwds.get("z_height", 0.0)
        self.curvature = kwds.get("curvature", 100.0)

    def _rpx_fill_face(self, rec, b):
        rec["type"] = _abi.FACE_SHAPED_SPHERICAL
        rec["p"][:2] = (self.curvature, self.z_height)
        self._fill_shape(rec, b)


class ExtrudedPlanarFace(Face):
    """Vertical plane through the segment (x1,y1)-(x2,y2), z1 < z < z2
    (cfaces.pyx:611-709)."""

    def __init__(self, **kwds):
        super().__init__(**_base_kwds(kwds))
        self._x1 = float(kwds.get('x1', 0))
        self._y1 = float(kwds.get('y1', 0))
        self._x2 = float(kwds.get('x2', 0))
        self._y2 = float(kwds.get('y2', 0))
        self.z1 = kwds.get('z1', 0)
        self.z2 = kwds.get('z2', 0)

    x1 = property(lambda s: s._x1, lambda s, v: setattr(s, "_x1", float(v)))
    y1 = property(lambda s: s._y1, lambda s, v: setattr(s, "_y1", float(v)))
    x2 = property(lambda s: s._x2, lambda s, v: setattr(s, "_x2", float(v)))
    y2 = property(lambda s: s._y2, lambda s, v: setattr(s, "_y2", float(v)))

    @property
    def normal(self):
        # calc_normal (cfaces.pyx:657-663)
        return norm((self._y1 - self._y2, self._x2 - self._x1, 0.0))

    def _rpx_fill_face(self, rec, b):
        rec["type"] = _abi.FACE_EXTRUDED_PLANAR
        n = self.normal
        rec["p"][:9] = (self._x1, self._y1, self._x2, self._y2, self.z1, self.z2, n[0], n[1], n[2])

    def compute_tangent(self, p):
        return (0.0, 0.0, 1.0)


class ConicRevolutionFace(ShapedFace):
    """Conic surface of revolution (cfaces.pyx:1751-1837)."""
    params = []

    def __init__(self, **kwds):
        super().__init__(**kwds)
        self.z_height = kwds.get('z_height', 0.0)
        self.conic_const = kwds.get('conic_const', 0.0)
        self.curvature = kwds.get('curvature', 10.0)

    def _rpx_fill_face(self, rec, b):
        rec["type"] = _abi.FACE_CONIC
        rec["p"][:3] = (self.curvature, self.z_height, self.conic_const)
        self._fill_shape(rec, b)


class AsphericFace(ShapedFace):
    """Even asphere, Newton root solve seeded by the conic (cfaces.pyx:1885-2025)."""
    params = []

    def __init__(self, **kwds):
        super().__init__(**kwds)
        self.z_height = kwds.get('z_height', 0.0)
        self.conic_const = kwds.get('conic_const', 0.0)
        self.curvature = kwds.get('curvature', 25.0)
        for k in ("A4", "A6", "A8", "A10", "A12", "A14", "A16"):
            setattr(self, k, kwds.get(k, 0.0))
        self.atol = kwds.get("atol", 1.0e-8)

    def _sag_band(self):
        """Half-width of the z band holding the surface over the aperture and
        a bound on |dsag/dr| there, or (-1, 0) when either is unbounded."""
        from .cshapes import bound_radius
        rb = bound_radius(self.shape)
        R = -float(self.curvature)
        beta = 1.0 + float(self.conic_const)
        A = [self.A4, self.A6, self.A8, self.A10, self.A12, self.A14, self.A16]
        if not math.isfinite(rb) or R == 0.0:
            return -1.0, 0.0
        arg = 1.0 - max(beta, 0.0) * rb * rb / (R * R)
        if arg <= 1e-6:
            return -1.0, 0.0
        band = rb * rb / abs(R) + sum(abs(a) * rb ** (2 * k + 4) for k, a in enumerate(A))
        slope = rb / (abs(R) * math.sqrt(arg)) + sum((2 * k + 4) * abs(a) * rb ** (2 * k + 3) for k, a in enumerate(A))
        if not (math.isfinite(band) and math.isfinite(slope)):
            return -1.0, 0.0
        return band, slope

    def _rpx_fill_face(self, rec, b):
        rec["type"] = _abi.FACE_ASPHERIC
        R = -float(self.curvature)
        band, slope = self._sag_band()
        # R**3 through the C library's pow, as the reference evaluates it
        rec["p"][:14] = (self.curvature, self.z_height, self.conic_const, self.A4, self.A6, self.A8,
                         self.A10, self.A12, self.A14, self.A16, self.atol, R ** 3, band, slope)
        self._fill_shape(rec, b)


class DistortionFace(ShapedFace):
    """A ShapedFace with a Distortion sag added (cfaces.pyx:2323-2438)."""

    def __init__(self, **kwds):
        face = kwds.get("base_face")
        kw = dict(kwds)
        kw.setdefault("shape", face.shape)
        super().__init__(**kw)
        self.base_face = face
        self.distortion = kwds.get("distortion")
        self.accuracy = kwds.get("accuracy", 1e-6)

    def _rpx_fill_face(self, rec, b):
        if isinstance(self.base_face, DistortionFace):
            from ._scene import UnsupportedOnDevice
            raise UnsupportedOnDevice("nested DistortionFace")
        rec["type"] = _abi.FACE_DISTORTION
        rec["p"][0] = self.accuracy
        self._fill_shape(rec, b)
        rec["aux0"] = b.extra_face(self.base_face)
        rec["pad"] = 1   # aux0 is relative to the extra-face block until finish()
        rec["aux1"] = b.distortion_index(self.distortion)
