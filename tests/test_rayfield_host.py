"""The oracle's RayField restatements (core/fields.py:51-111) on the CPU,
pinned by the reference's own tests (test_fields.py:14-70, with the source
built the current way: spacing = radius / resolution) and closed forms."""
import numpy as np

import oracle
from oracle import fields as ofields
from oracle import sources as osrc


def _field(spacing, radius=10.0, origin=(0, 0, -50)):
    rays, nb = osrc.hexagonal_ray_field(origin, (0.0, 0.0, 1.0), radius, radius / spacing, 2.0, (1.0, 0.0, 0.0),
                                        1.0 + 0j, 0j)
    return rays, nb


def test_project_to_sphere_reference_case():
    """test_fields.py:14-29: the field 50 mm before a sphere of radius 20 at
    the origin projects back to z = -20 on axis."""
    rays, _ = _field(2.0)
    out = ofields.project_to_sphere(rays.copy(), centre=(0, 0, 0), radius=20.0)
    assert len(out) == len(rays)
    assert abs(out["origin"][:, 2].min() - (-20.0)) < 1e-7


def test_project_to_sphere_closed_form():
    rng = np.random.default_rng(8)
    rays, _ = _field(1.0)
    rays["direction"] = rng.normal(size=(len(rays), 3)) * 0.05 + np.array([0, 0, 1.0])
    rays["direction"] /= np.linalg.norm(rays["direction"], axis=1)[:, None]
    rays["refractive_index"] = 1.33 + 0j
    c = np.array([0.5, -0.3, 2.0])
    out = ofields.project_to_sphere(rays.copy(), centre=c, radius=6.0)
    assert 0 < len(out) < len(rays)          # the outer rays miss the sphere
    np.testing.assert_allclose(np.linalg.norm(out["origin"] - c, axis=1), 6.0, rtol=1e-13)
    keep = np.isin(rays["ray_ident"], out["ray_ident"])
    src = rays[keep]
    moved = np.einsum("ij,ij->i", out["origin"] - src["origin"], src["direction"])
    np.testing.assert_allclose(out["accumulated_path"], 1.33 * moved, rtol=1e-12, atol=1e-12)
    # the nearer-to-minus-infinity root: the sphere's far side is ahead
    assert np.all(moved > 0)


def test_evaluate_neighbours_reference_case():
    """test_fields.py:32-48: 19 rays at spacing 4, 7 of them fully
    surrounded, neighbours at equal distance, collimated (dx = dy = 0)."""
    rays, nb = _field(4.0)
    assert len(rays) == 19
    sel, x, y, dx, dy = ofields.evaluate_neighbours(rays, nb)
    assert len(sel) == 7 and x.shape == (7, 6)
    off = x ** 2 + y ** 2
    assert np.allclose(off, off.mean())
    assert np.allclose(dx, 0) and np.allclose(dy, 0)


def test_modes_of_a_hexagonal_ring():
    """Six neighbours at radius s on a plane wave: x^2/s^2 + y^2/s^2 = 1
    exactly, so the fit gives A = C = i*blending/s^2, B = 0, no curvature."""
    rays, nb = _field(4.0)
    sel, x, y, dx, dy = ofields.evaluate_neighbours(rays, nb)
    m = oracle.evaluate_modes(x, y, dx, dy, blending=1.5)
    np.testing.assert_allclose(m[:, 0].imag, 1.5 / 16.0, rtol=1e-12)
    np.testing.assert_allclose(m[:, 2].imag, 1.5 / 16.0, rtol=1e-12)
    assert np.abs(m[:, 1]).max() < 1e-15 and np.abs(m.real).max() < 1e-15


def _lstsq_modes(x, y, dx, dy, blending):
    """The least-squares problems of core/fields.py:139-175, solved
    independently with numpy.linalg.lstsq one gausslet at a time: widths
    from [x^2, 2xy, y^2] c = 1, curvatures from the stacked rows
    [x, y, 0] r = dx and [0, x, y] r = dy."""
    out = np.zeros((x.shape[0], 3), np.complex128)
    for i in range(x.shape[0]):
        M = np.column_stack([x[i] ** 2, 2 * x[i] * y[i], y[i] ** 2])
        im = np.linalg.lstsq(M, np.ones(x.shape[1]), rcond=None)[0]
        z = np.zeros_like(x[i])
        R = np.vstack([np.column_stack([x[i], y[i], z]), np.column_stack([z, x[i], y[i]])])
        re = np.linalg.lstsq(R, np.concatenate([dx[i], dy[i]]), rcond=None)[0]
        out[i] = blending * 1j * im + re
    return out


def test_evaluate_modes_rotated_hexagon_reference_case():
    """test_fields.py:71-85: an elliptical hexagon (2:1) turned by 20 degrees
    with slopes dx = -x/10 + y/50, dy = y/10 + x/50 (taken at the unrotated
    points, as the reference test does).  The closed-form fit
    (cfields.pyx:156-228) must equal the least-squares solution of the
    reference's Python fit in both parts: the real (curvature) part is
    non-zero here, so a sign or index error in it fails.

    One difference is fixed by the algebra of the closed form itself: its
    cross term is the negative of the least-squares B, in both the real and
    the imaginary part (Cramer's rule on the same normal equations gives
    +B where cfields.pyx:191 and :215 carry a leading minus).  The
    reference's own comparison (test_fields.py:71-85) never runs -- it
    imports evaluate_modes from raypier.fields, which has none -- and its
    field path calls the closed form (core/fields.py:192), so the oracle and
    the device keep the closed form's sign: A and C equal the least-squares
    values and B equals minus the least-squares value."""
    angles = np.array([np.linspace(0, np.pi * 2, 7)[:6]])
    x = np.sin(angles) * 2.0
    y = np.cos(angles)
    rot = 20.0 * np.pi / 180.0
    x2 = np.cos(rot) * x - np.sin(rot) * y
    y2 = np.cos(rot) * y + np.sin(rot) * x
    dx = -x / 10.0 + y / 50.0
    dy = y / 10.0 + x / 50.0
    for blending in (1.0, 1.7):
        got = oracle.evaluate_modes(x2, y2, dx, dy, blending)
        want = _lstsq_modes(x2, y2, dx, dy, blending) * np.array([1.0, -1.0, 1.0])
        assert np.abs(want.real).min() > 0.01 and np.abs(want.imag).min() > 0.1
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)


def test_evaluate_modes_random_against_lstsq():
    rng = np.random.default_rng(31)
    x, y = rng.normal(size=(200, 6)), rng.normal(size=(200, 6))
    dx, dy = 0.1 * rng.normal(size=(200, 6)), 0.1 * rng.normal(size=(200, 6))
    np.testing.assert_allclose(oracle.evaluate_modes(x, y, dx, dy, 1.3),
                               _lstsq_modes(x, y, dx, dy, 1.3) * np.array([1.0, -1.0, 1.0]), rtol=1e-9, atol=1e-12)


def test_neighbours_gc_known_gausslet():
    """evaluate_neighbours_gc (core/fields.py:114-136) on a gausslet built
    in a rotated frame: base direction R z, E vector R x, parabasal ray j
    starting at R (a_j, b_j, 0) from the base origin and pointing along
    R normalize(p_j, q_j, 1).  With H = E x d = R (0, -1, 0) the expected
    values are x = a_j, y = -b_j, dx = p_j, dy = -q_j, which fixes the sign
    convention of H."""
    from raypier_optics_amd._abi import gausslet_dtype
    rng = np.random.default_rng(5)
    ax = rng.normal(size=3)
    ax /= np.linalg.norm(ax)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    R = np.eye(3) + np.sin(1.1) * K + (1 - np.cos(1.1)) * K @ K
    a, b = rng.normal(size=6), rng.normal(size=6)
    p, q = 0.01 * rng.normal(size=6), 0.01 * rng.normal(size=6)
    g = np.zeros(1, gausslet_dtype)
    o = np.array([1.0, -2.0, 0.5])
    g["base_ray"]["origin"] = o
    g["base_ray"]["direction"] = R @ np.array([0.0, 0.0, 1.0])
    g["base_ray"]["E_vector"] = R @ np.array([1.0, 0.0, 0.0])
    for j in range(6):
        g["para_rays"][0, j]["origin"] = o + R @ np.array([a[j], b[j], 0.0])
        d = np.array([p[j], q[j], 1.0])
        g["para_rays"][0, j]["direction"] = R @ (d / np.linalg.norm(d))
    X, Y, DX, DY = oracle.neighbours_gc(g)
    np.testing.assert_allclose(X[0], a, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(Y[0], -b, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(DX[0], p, rtol=1e-10, atol=1e-15)
    np.testing.assert_allclose(DY[0], -q, rtol=1e-10, atol=1e-15)


def test_host_evaluate_neighbours_gc_matches_pinned_oracle():
    """raypier_optics_amd.fields.evaluate_neighbours_gc (numpy, host) against
    oracle.neighbours_gc, itself pinned above by the rotated-frame gausslet."""
    from raypier_optics_amd import fields
    from raypier_optics_amd._abi import gausslet_dtype
    rng = np.random.default_rng(12)
    g = np.zeros(300, gausslet_dtype)
    d = rng.normal(size=(300, 3))
    d[:, 2] = np.abs(d[:, 2]) + 1.0
    d /= np.linalg.norm(d, axis=1)[:, None]
    e = np.cross(d, rng.normal(size=(300, 3)))
    e /= np.linalg.norm(e, axis=1)[:, None]
    g["base_ray"]["direction"], g["base_ray"]["E_vector"] = d, e
    g["base_ray"]["origin"] = rng.normal(size=(300, 3))
    g["para_rays"]["origin"] = g["base_ray"]["origin"][:, None, :] + 0.01 * rng.normal(size=(300, 6, 3))
    pd = d[:, None, :] + 0.01 * rng.normal(size=(300, 6, 3))
    g["para_rays"]["direction"] = pd / np.linalg.norm(pd, axis=2)[..., None]
    _, x, y, dx, dy = fields.evaluate_neighbours_gc(g)
    for got, want in zip((x, y, dx, dy), oracle.neighbours_gc(g)):
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15)
