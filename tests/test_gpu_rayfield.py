"""GPU parity of the RayField E-field path (rpx_project_to_sphere,
rpx_ray_field_modes + rpx_sum_gaussian_modes behind
fields.project_to_sphere / evaluate_neighbours / eval_Efield_from_rays)
against the oracle's restatement of core/fields.py:51-111, 196-218.

Bar: projected and selected rays, neighbour coordinates and slopes
bit-exact; modes to 1e-12 relative (the closed-form 3x3 solve, as the
gausslet mode test); the field to 1e-9 of its largest component (the field
sum's transcendentals differ between device and host libraries).
"""
import numpy as np
import pytest

import oracle
from oracle import fields as ofields
from raypier_optics_amd import workloads, fields
from raypier_optics_amd._lib import RpxError
from raypier_optics_amd.sources import HexagonalRayFieldSource
from raypier_optics_amd.tracer import trace_rays

pytestmark = pytest.mark.gpu


def _traced_field(n_gen=2):
    """A hexagonal ray field through the C1 singlet: the rays leaving the
    back face carry a converging wavefront and propagated neighbours."""
    w = workloads.singlet_c1(n_rays=10)
    src = HexagonalRayFieldSource(origin=(0, 0, -10), direction=(0, 0, 1), radius=8.0, resolution=25.0,
                                  wavelength=0.55)
    traced, _ = trace_rays(src.input_rays, w.face_lists, recursion_limit=4, max_length=100.0)
    return traced[n_gen], w


def test_project_to_sphere(gpu):
    rc, w = _traced_field()
    rays = rc.copy_as_array()
    for centre, radius in (((0.0, 0.0, 52.0), 30.0), ((0.5, -0.2, 40.0), 8.0)):
        got = fields.project_to_sphere(rays, centre, radius)
        ref = ofields.project_to_sphere(rays.copy(), centre, radius)
        assert len(ref) > 100
        assert got.tobytes() == ref.tobytes()


def test_evaluate_neighbours(gpu):
    rc, w = _traced_field()
    rays, nb = rc.copy_as_array(), rc.neighbours
    got = fields.evaluate_neighbours(rays, nb)
    ref = ofields.evaluate_neighbours(rays, nb)
    assert len(ref[0]) > 500
    assert got[0].tobytes() == ref[0].tobytes()
    for g, r, name in zip(got[1:], ref[1:], ("x", "y", "dx", "dy")):
        assert g.tobytes() == r.tobytes(), name


@pytest.mark.parametrize("pupil", [0.0, 30.0])
def test_eval_efield_from_rays(gpu, pupil):
    rc, w = _traced_field()
    rays, nb = rc.copy_as_array(), rc.neighbours
    focus = w.info["detector_z_mm"]
    g = np.linspace(-0.02, 0.02, 24)
    X, Y = np.meshgrid(g, g)
    pts = np.stack([X.ravel(), Y.ravel(), np.full(X.size, focus - 0.5)], axis=1)
    kw = dict(blending=1.0, time_ps=0.0)
    if pupil:
        kw.update(exit_pupil_offset=pupil, exit_pupil_centre=(0.0, 0.0, focus))
    got = fields.eval_Efield_from_rays(rc, pts, rc.wavelengths, **kw)
    ref = ofields.eval_Efield_from_rays(rays, nb, pts, rc.wavelengths, **kw)
    scale = np.abs(ref).max()
    assert scale > 0
    assert np.abs(got - ref).max() <= 1e-9 * scale
    # the modes themselves
    sel, x, y, dx, dy = ofields.evaluate_neighbours(ofields.project_to_sphere(rays.copy(), (0, 0, focus), pupil)
                                                    if pupil else rays, nb)
    ref_modes = oracle.evaluate_modes(x, y, dx, dy, blending=1.0)
    dev = rc._ensure_device()
    if pupil:
        dev, _ = fields._project_device(dev, rc.n_rays, (0.0, 0.0, focus), pupil)
    _, k, modes, _ = fields._ray_modes_device(dev, rc.n_rays, nb, 1.0)
    m = modes.cpu().numpy().view(np.complex128)
    assert k == len(ref_modes)
    assert np.allclose(m, ref_modes, rtol=1e-12, atol=0)


def test_eval_efield_errors(gpu):
    rc, w = _traced_field()
    # a sphere that some rays miss leaves the neighbour table misaligned
    with pytest.raises(IndexError):
        fields.eval_Efield_from_rays(rc, np.zeros((4, 3)), rc.wavelengths, exit_pupil_offset=2.0,
                                     exit_pupil_centre=(0.0, 0.0, 60.0))
    # a neighbour index past the rays, in a fully surrounded row (rows with
    # a -1 are skipped before their indices are used, as in the reference)
    bad = rc.neighbours.copy()
    row = int(np.nonzero((bad >= 0).all(axis=1))[0][0])
    bad[row, 2] = rc.n_rays + 5
    with pytest.raises(RpxError):
        fields.evaluate_neighbours(rc.copy_as_array(), bad)
    edge = rc.neighbours.copy()
    row = int(np.nonzero(~(edge >= 0).all(axis=1))[0][0])
    edge[row, int(np.nonzero(edge[row] >= 0)[0][0])] = rc.n_rays + 5
    assert len(fields.evaluate_neighbours(rc.copy_as_array(), edge)[0]) == \
        len(ofields.evaluate_neighbours(rc.copy_as_array(), rc.neighbours)[0])
