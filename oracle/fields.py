"""Oracle restatements of the RayField E-field path, in numpy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  project_to_sphere and
evaluate_neighbours follow core/fields.py:51-111 statement by statement;
eval_Efield_from_rays (core/fields.py:196-218) chains them with the oracle's
C mode fit and field sum.
"""
import numpy

import oracle


def project_to_sphere(rays, centre=(0, 0, 0), radius=10.0):
    centre = numpy.asarray(centre).reshape(1, 3)
    origin = rays['origin'] - centre
    direction = rays['direction']
    c = (origin ** 2).sum(axis=1) - (radius ** 2)
    b = 2 * (direction * origin).sum(axis=1)
    d = b ** 2 - 4 * c
    selector = (d >= 0.0)
    d = numpy.sqrt(d[selector])
    b = b[selector]
    root1 = (-b + d) / 2
    root2 = (-b - d) / 2
    alpha = numpy.column_stack((root1, root2)).min(axis=1)
    rays = rays[selector]
    rays['origin'] += alpha[:, None] * direction[selector]
    rays['accumulated_path'] += alpha * rays['refractive_index'].real
    return rays


def evaluate_neighbours(rays, neighbours_idx):
    mask = (neighbours_idx >= 0).all(axis=1)
    nb = neighbours_idx[mask, :]
    n_origin = rays['origin'][nb, :]
    n_direction = rays['direction'][nb, :]
    origin = rays['origin'][mask, None, :]
    direction = rays['direction'][mask, None, :]
    E = rays['E_vector'][mask, None, :]
    H = numpy.cross(E, direction)
    offsets = n_origin - origin
    alpha = -(offsets * direction).sum(axis=-1) / (n_direction * direction).sum(axis=-1)
    projected = (offsets + alpha[:, :, None] * n_direction)
    x = (projected * E).sum(axis=-1)
    y = (projected * H).sum(axis=-1)
    dz = (n_direction * direction).sum(axis=-1)
    dx = (n_direction * E).sum(axis=-1) / dz
    dy = (n_direction * H).sum(axis=-1) / dz
    return (rays[mask], x, y, dx, dy)


def eval_Efield_from_rays(rays, neighbours_idx, points, wavelengths, blending=1.0, time_ps=0.0,
                          exit_pupil_offset=0.0, exit_pupil_centre=(0.0, 0.0, 0.0)):
    radius = exit_pupil_offset
    if radius:
        projected = project_to_sphere(rays, exit_pupil_centre, radius)
    else:
        projected = rays
    sel, x, y, dx, dy = evaluate_neighbours(projected, neighbours_idx)
    modes = oracle.evaluate_modes(x, y, dx, dy, blending=blending)
    return oracle.sum_gaussian_modes(sel, modes, wavelengths, points, time_ps)
