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
