"""E-field evaluation from gausslets (drop-in for raypier.core.fields)."""
import numpy

from .cfields import sum_gaussian_modes, evaluate_modes as evaluate_modes_c, gausslet_modes
from .ctracer import RayCollection, GaussletCollection


def evaluate_neighbours_gc(gausslets):
    """Parabasal offsets and slopes in the (E, E x d) frame of each base ray
    (core/fields.py:114-136); host numpy on a gausslet_dtype array."""
    ga = gausslets
    origin = ga['base_ray']['origin'][:, None, :]
    direction = ga['base_ray']['direction'][:, None, :]
    n_direction = ga['para_rays']['direction']
    E = ga['base_ray']['E_vector'][:, None, :]
    H = numpy.cross(E, direction)
    offsets = ga['para_rays']['origin'] - origin
    x = (offsets * E).sum(axis=-1)
    y = (offsets * H).sum(axis=-1)
    dz = (n_direction * direction).sum(axis=-1)
    dx = (n_direction * E).sum(axis=-1) / dz
    dy = (n_direction * H).sum(axis=-1) / dz
    return (ga['base_ray'], x, y, dx, dy)


def Gamma(z, A, B, C):
    detG0 = (A * C) - (B * B)
    denom = (1 + (z * (A + C)) + (z * z) * detG0)
    return (A + z * detG0) / denom, B / denom, (C + z * detG0) / denom


def _base_collection(gc):
    """The base rays of a GaussletCollection as a RayCollection sharing its
    device buffer (the SoA ray rows are identical)."""
    dev = gc._ensure_device()
    rc = RayCollection.from_device(dev, gc.n_rays)
    rc.wavelengths = gc.wavelengths
    return rc


class EFieldSummation(object):
    """Mode coefficients computed once, field evaluated at many point sets
    (core/fields.py:221-249)."""

    def __init__(self, gausslet_collection, wavelengths=None, blending=1.0):
        if wavelengths is None:
            wavelengths = numpy.asarray(gausslet_collection.wavelengths)
        if wavelengths is None:
            raise ValueError("No wavelengths supplied")
        self.wavelengths = wavelengths
        self.gc = gausslet_collection
        self.modes = gausslet_modes(gausslet_collection, blending)
        self.base_rays = _base_collection(gausslet_collection)

    def evaluate(self, points, time_ps=0.0):
        points = numpy.ascontiguousarray(points)
        shape = points.shape
        E = sum_gaussian_modes(self.base_rays, self.modes, self.wavelengths, points.reshape(-1, 3), time_ps)
        return E.reshape(shape)


def eval_Efield_from_gausslets(gausslet_collection, points, wavelengths=None, blending=1.0, time_ps=0.0,
                               **kwds):
    """Vector E-field at `points` from a GaussletCollection
    (core/fields.py:252-276); complex128 (N, 3)."""
    if wavelengths is None:
        wavelengths = numpy.asarray(gausslet_collection.wavelengths)
    modes = gausslet_modes(gausslet_collection, blending)
    rays = _base_collection(gausslet_collection)
    return sum_gaussian_modes(rays, modes, wavelengths, points, time_ps, precision=kwds.get("precision", 0))
