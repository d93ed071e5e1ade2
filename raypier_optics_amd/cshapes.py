"""2-D aperture shapes (drop-in for raypier.core.cshapes).

A shape tree compiles to a postfix program of the scene table; the predicate
itself runs on the GPU (cshapes.pyx:34-168 semantics).
"""
import math

import numpy as np

from . import _abi
from .ctracer import Shape


def bound_radius(shape):
    """Radius of a disk about the origin containing every inside point of
    the shape (infinite when unbounded)."""
    f = getattr(shape, "_rpx_bound_radius", None)
    return f() if f is not None else float("inf")


class LogicalOpShape(Shape):
    def __and__(self, other):
        return BooleanAND(self, other)

    def __or__(self, other):
        return BooleanOR(self, other)

    def __xor__(self, other):
        return BooleanXOR(self, other)

    def __invert__(self):
        return InvertShape(self)


class InvertShape(LogicalOpShape):
    """cshapes.pyx:34-42"""

    def __init__(self, shape):
        self.shape = shape

    def _rpx_bound_radius(self):
        return float("inf")

    def _rpx_children(self):
        return [self.shape]

    def _rpx_shape_ops(self, rec, builder):
        rec["op"] = _abi.SHAPE_NOT


class BooleanShape(LogicalOpShape):
    _op = None

    def __init__(self, shape1, shape2):
        self.shape1 = shape1
        self.shape2 = shape2

    def _rpx_children(self):
        return [self.shape1, self.shape2]

    def _rpx_shape_ops(self, rec, builder):
        rec["op"] = self._op


class BooleanAND(BooleanShape):
    _op = _abi.SHAPE_AND

    def _rpx_bound_radius(self):
        return min(bound_radius(self.shape1), bound_radius(self.shape2))


class BooleanOR(BooleanShape):
    _op = _abi.SHAPE_OR

    def _rpx_bound_radius(self):
        return max(bound_radius(self.shape1), bound_radius(self.shape2))


class BooleanXOR(BooleanShape):
    _op = _abi.SHAPE_XOR

    def _rpx_bound_radius(self):
        return max(bound_radius(self.shape1), bound_radius(self.shape2))


class BasicShape(LogicalOpShape):
    def __init__(self, **kwds):
        if "centre" in kwds:
            self.centre = kwds["centre"]
        else:
            self.centre_x = kwds.get("centre_x", 0.0)
            self.centre_y = kwds.get("centre_y", 0.0)

    @property
    def centre(self):
        return (self.centre_x, self.centre_y)

    @centre.setter
    def centre(self, v):
        self.centre_x = float(v[0])
        self.centre_y = float(v[1])


class CircleShape(BasicShape):
    """cshapes.pyx:102-116: inside iff dx^2 + dy^2 < r^2."""

    def __init__(self, **kwds):
        super().__init__(**kwds)
        self.radius = kwds.get("radius", 1.0)

    def _rpx_bound_radius(self):
        return math.hypot(self.centre_x, self.centre_y) + abs(self.radius)

    def _rpx_shape_ops(self, rec, builder):
        rec["op"] = _abi.SHAPE_CIRCLE
        rec["p"][:3] = (self.centre_x, self.centre_y, self.radius)


class RectangleShape(BasicShape):
    """cshapes.pyx:119-136: inside iff 2|dx| < width and 2|dy| < height."""

    def __init__(self, **kwds):
        super().__init__(**kwds)
        self.width = kwds.get("width", 5.0)
        self.height = kwds.get("height", 7.0)

    def _rpx_bound_radius(self):
        return math.hypot(self.centre_x, self.centre_y) + 0.5 * math.hypot(self.width, self.height)

    def _rpx_shape_ops(self, rec, builder):
        rec["op"] = _abi.SHAPE_RECT
        rec["p"][:4] = (self.centre_x, self.centre_y, self.width, self.height)


class PolygonShape(BasicShape):
    """cshapes.pyx:139-168: even-odd crossing test."""

    def __init__(self, **kwds):
        super().__init__(**kwds)
        self._coordinates = np.zeros((0, 2))
        if "coordinates" in kwds:
            self.coordinates = kwds["coordinates"]

    @property
    def coordinates(self):
        return np.asarray(self._coordinates)

    @coordinates.setter
    def coordinates(self, val):
        self._coordinates = np.ascontiguousarray(val, dtype=np.float64).reshape(-1, 2)

    def _rpx_bound_radius(self):
        if len(self._coordinates) == 0:
            return float("inf")
        return float(np.hypot(self._coordinates[:, 0], self._coordinates[:, 1]).max())

    def _rpx_shape_ops(self, rec, builder):
        off, n = builder.add_vertices(self._coordinates)
        rec["op"] = _abi.SHAPE_POLYGON
        rec["a"] = off
        rec["b"] = n
