"""Beam-decomposition pieces on the CPU: the oracle's restatements of
calc_mode_curvature, apply_mode_curvature and build_interaction_matrix
(cfields.pyx:231-424) pinned by closed forms, and the host-side helpers
(unwrap2d, make_hexagonal_grid) against the oracle's copies and closed forms.
"""
import math

import numpy as np
import pytest

import oracle
from oracle import gausslets as og
from raypier_optics_amd import workloads
from raypier_optics_amd.ctracer import GaussletCollection
from raypier_optics_amd.gausslets import make_hexagonal_grid, next_power_of_two
from raypier_optics_amd.unwrap2d import unwrap2d


def test_mode_curvature_plane_wave():
    """Tilted plane wavefront (constant slopes, no second derivatives): no
    curvature, z' along the wavefront normal, an orthonormal basis."""
    rng = np.random.default_rng(1)
    dx, dy = rng.uniform(-0.2, 0.2, 50), rng.uniform(-0.2, 0.2, 50)
    z0 = np.zeros(50)
    A, B, C, x, y, z = oracle.calc_mode_curvature(z0, z0, dx, dy, z0, z0, z0)
    assert np.all(A == 0) and np.all(B == 0) and np.all(C == 0)
    n = np.stack([-dx, -dy, np.ones(50)], axis=1)
    np.testing.assert_allclose(z, n / np.linalg.norm(n, axis=1)[:, None], rtol=1e-15)
    for u, v in ((x, y), (y, z), (x, z)):
        np.testing.assert_allclose(np.einsum("ij,ij->i", u, v), 0.0, atol=1e-15)
    for u in (x, y, z):
        np.testing.assert_allclose(np.linalg.norm(u, axis=1), 1.0, rtol=1e-15)


def test_mode_curvature_sphere_limit_and_on_axis():
    """z = (x^2 + y^2) / 2R near the axis: A = C -> -1/R, B -> 0.  Exactly on
    the axis the reference's basis is undefined (cross((0,0,1), z') = 0,
    normalised): NaN, as there."""
    R = 50.0
    x0 = np.array([1e-3, 0.0, 1e-3 / math.sqrt(2)])
    y0 = np.array([0.0, 1e-3, 1e-3 / math.sqrt(2)])
    one = np.full(3, 1.0 / R)
    A, B, C, *_ = oracle.calc_mode_curvature(x0, y0, x0 / R, y0 / R, one, one, np.zeros(3))
    np.testing.assert_allclose(A, -1.0 / R, rtol=1e-6)
    np.testing.assert_allclose(C, -1.0 / R, rtol=1e-6)
    assert np.abs(B).max() < 1e-9
    A0, *_ = oracle.calc_mode_curvature([0.0], [0.0], [0.0], [0.0], [1 / R], [1 / R], [0.0])
    assert np.isnan(A0[0])


def _gausslets(n, seed):
    rng = np.random.default_rng(seed)
    r = workloads.base_rays(n)
    d = rng.normal(size=(n, 3)) * 0.1 + np.array([0, 0, 1.0])
    r["direction"] = d / np.linalg.norm(d, axis=1)[:, None]
    e = np.cross(r["direction"], np.array([0.0, 1.0, 0.0]))
    r["E_vector"] = e / np.linalg.norm(e, axis=1)[:, None]
    r["origin"] = rng.uniform(-2, 2, (n, 3))
    r["wavelength_idx"] = 0
    g = GaussletCollection.from_rays(r).copy_as_array()
    return oracle.config_parabasal_rays(g, np.array([0.8]), 0.05, 0.0)


def test_apply_mode_curvature_first_order():
    """Zero coefficients leave the (unit) directions as they were; small
    ones tilt each parabasal ray by -(A px + B py) x - (C py + B px) y to
    first order (px, py: its offset in the base ray's (E, E x d) frame)."""
    g = _gausslets(40, 2)
    z = np.zeros(40)
    same = oracle.apply_mode_curvature(g, z, z, z)
    np.testing.assert_allclose(same["para_rays"]["direction"], g["para_rays"]["direction"], atol=2e-16)
    A, B, C = np.full(40, 1e-4), np.full(40, -3e-5), np.full(40, 2e-4)
    out = oracle.apply_mode_curvature(g, A, B, C)
    b = g["base_ray"]
    xx = b["E_vector"]
    yy = np.cross(xx, b["direction"])
    for j in range(6):
        r = g["para_rays"][:, j]["origin"] - b["origin"]
        px, py = np.einsum("ij,ij->i", r, xx), np.einsum("ij,ij->i", r, yy)
        shift = xx * (-A * px - B * py)[:, None] + yy * (-C * py - B * px)[:, None]
        expect = g["para_rays"][:, j]["direction"] + shift
        expect /= np.linalg.norm(expect, axis=1)[:, None]
        np.testing.assert_allclose(out["para_rays"][:, j]["direction"], expect, atol=1e-15)


def test_interaction_matrix_closed_forms():
    """Pairs are exactly those within max_spacing (brute force); a flat mode
    (no curvature, basis = lab axes) at its own origin gives 1 -- the
    docstring's E_ii = 1 -- and at distance r has magnitude
    exp(-(blending * r / spacing)^2)."""
    rx, ry = make_hexagonal_grid(1.0, spacing=0.25)
    tx, ty = make_hexagonal_grid(1.0, spacing=0.125)
    n = len(rx)
    zero = np.zeros(n)
    eye = np.tile(np.eye(3), (n, 1, 1))
    x, y, z = eye[:, 0, :], eye[:, 1, :], eye[:, 2, :]
    wl, spacing, max_spacing, blending = 0.8, 0.25, 0.4, 1.3
    data, (i, j) = oracle.build_interaction_matrix(rx, ry, tx, ty, zero, zero, zero, x, y, z, wl, spacing,
                                                   max_spacing, blending)
    d2 = (tx[None, :] - rx[:, None]) ** 2 + (ty[None, :] - ry[:, None]) ** 2
    bi, bj = np.nonzero(d2 <= max_spacing ** 2)
    assert np.array_equal(i, bi) and np.array_equal(j, bj)
    r2 = d2[i, j]
    np.testing.assert_allclose(np.abs(data), np.exp(-(blending ** 2) * r2 / spacing ** 2), rtol=1e-12)
    same = r2 == 0.0
    assert same.sum() > 10
    np.testing.assert_allclose(data[same], 1.0 + 0j, atol=1e-15)


def test_unwrap2d():
    """A smooth phase surface wrapped into (-pi, pi] comes back whole, with
    the anchor pixel's value kept; residues vanish."""
    g = np.linspace(-1, 1, 40)
    X, Y = np.meshgrid(g, g)
    true = 10.0 * (X ** 2 + 0.5 * Y ** 2) + 2.0 * X      # steps below pi per pixel
    wrapped = np.angle(np.exp(1j * true))
    out, res = unwrap2d(wrapped, anchor=(20, 20))
    ref_out, ref_res = og.unwrap2d(wrapped, anchor=(20, 20))
    assert out.tobytes() == ref_out.tobytes() and res.tobytes() == ref_res.tobytes()
    assert abs(out[20, 20] - wrapped[20, 20]) < 1e-12
    k = np.round((out - true) / (2 * np.pi))
    assert np.all(k == k[0, 0])
    np.testing.assert_allclose(out - true, k * 2 * np.pi, atol=1e-9)
    assert np.abs(res).max() < 1e-9


@pytest.mark.parametrize("radius,spacing", [(10.0, 4.0), (1.0, 0.1), (3.3, 0.37)])
def test_hexagonal_grid_connectivity(radius, spacing):
    x, y, nb = make_hexagonal_grid(radius, spacing, connectivity=True)
    rx, ry, rnb = og.make_hexagonal_grid(radius, spacing, connectivity=True)
    assert x.tobytes() == rx.tobytes() and y.tobytes() == ry.tobytes()
    assert np.array_equal(nb, rnb)
    for k in range(len(x)):
        for q in nb[k]:
            if q >= 0:
                assert abs(math.hypot(x[q] - x[k], y[q] - y[k]) - spacing) < 1e-9 * spacing


def test_next_power_of_two():
    assert [next_power_of_two(n) for n in (1, 2, 3, 64, 65, 1000)] == [1, 2, 4, 64, 128, 1024]
