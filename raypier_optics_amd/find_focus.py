"""Least-squares focus of a bundle of rays (drop-in for
raypier.core.find_focus, core/find_focus.py:10-52).

The focus is the point x minimising the (weighted) sum of squared distances
to the ray lines: sum_i w_i (n_i n_i^T - I)(x - o_i) = 0.  As in the
reference, n_i is the direction divided by its squared length (the same as
the unit direction for unit directions) and the weights are the rays'
normalised power.
"""
import numpy as np


def find_focus(ray_origins, ray_directions, weights=None):
    o = np.asarray(ray_origins, dtype=np.float64)
    d = np.asarray(ray_directions, dtype=np.float64)
    n = d / (d * d).sum(axis=1, keepdims=True)
    A = n[:, :, None] * n[:, None, :] - np.eye(3)[None, :, :]
    if weights is not None:
        A = A * np.asarray(weights, dtype=np.float64)[:, None, None]
    AA = A.sum(axis=0)
    b = np.einsum("nij,nj->i", A, o)
    return np.linalg.solve(AA, b)


def find_ray_focus(ray_collection):
    """Focus of a RayCollection or of a GaussletCollection's base rays,
    weighted by power |E1|^2 + |E2|^2 (core/find_focus.py:10-27)."""
    from .ctracer import GaussletCollection, RayCollection
    if isinstance(ray_collection, GaussletCollection):
        ray = ray_collection.copy_as_array()["base_ray"]
        E1, E2, origin, direction = ray["E1_amp"], ray["E2_amp"], ray["origin"], ray["direction"]
    elif isinstance(ray_collection, RayCollection):
        E1, E2 = ray_collection.E1_amp, ray_collection.E2_amp
        origin, direction = ray_collection.origin, ray_collection.direction
    else:
        raise TypeError("Expecting a RayCollection object")
    weights = E1.real ** 2 + E1.imag ** 2 + E2.real ** 2 + E2.imag ** 2
    weights = weights / weights.sum()
    return find_focus(origin, direction, weights=weights)
