"""Oracle restatements of the reference's lattice sources, in numpy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Each function follows the
reference's input_rays construction statement by statement and returns plain
ray_dtype / gausslet_dtype arrays:

  hexagonal_ray_field   sources.py:1035-1107  (HexagonalRayFieldSource)
  collimated_gausslets  gausslet_sources.py:386-457 (CollimatedGaussletSource)
  gaussian_point_source gausslet_sources.py:516-566 (GaussianPointSource)
  make_hexagonal_grid   core/gausslets.py:55-80 (without connectivity)
"""
import numpy

import oracle
from raypier_optics_amd._abi import ray_dtype, gausslet_dtype


def normaliseVector(a):
    """core/utils.py:5-9"""
    a = numpy.asarray(a)
    mag = numpy.sqrt((a ** 2).sum(axis=-1))[..., numpy.newaxis]
    return a / mag


def _from_rays(data):
    """GaussletCollection.from_rays (ctracer.pyx:1321-1345)."""
    g = numpy.zeros(len(data), gausslet_dtype)
    g["base_ray"] = data
    for j in range(6):
        p = g["para_rays"][:, j]
        p["origin"] = data["origin"]
        p["direction"] = data["direction"]
        p["normal"] = data["normal"]
        p["length"] = data["length"]
    return g


def hexagonal_ray_field(origin, direction, radius, resolution, gauss_width, E_vector, E1_amp, E2_amp):
    """(ray_dtype array, neighbours (n, 6) int32)."""
    origin = numpy.array(origin)
    direction = numpy.array(direction)
    spacing = radius / resolution
    max_axis = numpy.abs(direction).argmax()
    if max_axis == 0:
        v = numpy.array([0., 1., 0.])
    else:
        v = numpy.array([1., 0., 0.])
    d1 = numpy.cross(direction, v)
    d1 = normaliseVector(d1)
    d2 = numpy.cross(direction, d1)
    d2 = normaliseVector(d2)

    E_vector = numpy.cross(E_vector, direction)
    E_vector = numpy.cross(E_vector, direction)

    cosV = numpy.cos(numpy.pi * 30 / 180.)
    sinV = numpy.sin(numpy.pi * 30 / 180.)
    d2 = cosV * d2 + sinV * d1

    nsteps = int(radius / (spacing * numpy.cos(numpy.pi * 30 / 180.)))
    i = numpy.arange(-nsteps - 1, nsteps + 2)
    j = numpy.arange(-nsteps - 1, nsteps + 2)

    vi, vj = numpy.meshgrid(i, j)
    label = numpy.arange(vi.size).reshape(*vi.shape)
    neighbours = numpy.full((label.shape[0], label.shape[1], 6), -1, 'i')
    neighbours[1:, :, 3] = label[:-1, :]
    neighbours[:-1, :, 0] = label[1:, :]
    neighbours[:, 1:, 4] = label[:, :-1]
    neighbours[:, :-1, 1] = label[:, 1:]
    neighbours[1:, :-1, 2] = label[:-1, 1:]
    neighbours[:-1, 1:, 5] = label[1:, :-1]
    neighbours.shape = -1, 6

    vi.shape = -1
    vj.shape = -1

    ri = ((vi + sinV * vj) ** 2 + (cosV * vj ** 2))
    select = ri < (radius / spacing) ** 2

    xi = vi[select]
    yj = vj[select]

    backmap = numpy.full(vi.shape[0] + 1, -1, 'i')
    backmap[:-1][select] = numpy.arange(len(xi))
    selnb = backmap[neighbours]
    neighbours = selnb[select, :]

    offsets = xi[:, None] * d1 + yj[:, None] * d2
    offsets *= spacing

    gauss = numpy.exp(-(ri[select] * (spacing ** 2) / ((gauss_width) ** 2)))

    n_rays = offsets.shape[0]
    ray_data = numpy.zeros(n_rays, dtype=ray_dtype)
    ray_data['origin'] = offsets
    ray_data['origin'] += origin
    ray_data['direction'] = direction
    ray_data['wavelength_idx'] = 0
    ray_data['E_vector'] = [normaliseVector(E_vector)]
    ray_data['E1_amp'] = E1_amp * gauss
    ray_data['E2_amp'] = E2_amp * gauss
    ray_data['refractive_index'] = 1.0 + 0.0j
    ray_data['normal'] = [[0, 1, 0]]
    ray_data['ray_ident'] = numpy.arange(n_rays)
    return ray_data, neighbours


def collimated_gausslets(origin, direction, radius, resolution, beam_waist, blending, wavelength, beam_power,
                         E_vector, E1_amp, E2_amp):
    origin = numpy.array(origin)
    direction = numpy.array(direction)
    spacing = radius / resolution
    max_axis = numpy.abs(direction).argmax()
    if max_axis == 0:
        v = numpy.array([0., 1., 0.])
    else:
        v = numpy.array([1., 0., 0.])
    d1 = numpy.cross(direction, v)
    d1 = normaliseVector(d1)
    d2 = numpy.cross(direction, d1)
    d2 = normaliseVector(d2)

    E_vector = numpy.cross(E_vector, direction)
    E_vector = numpy.cross(E_vector, direction)

    cosV = numpy.cos(numpy.pi * 30 / 180.)
    sinV = numpy.sin(numpy.pi * 30 / 180.)
    d2 = cosV * d2 + sinV * d1

    nsteps = int(radius / (spacing * numpy.cos(numpy.pi * 30 / 180.)))
    i = numpy.arange(-nsteps - 1, nsteps + 2)
    j = numpy.arange(-nsteps - 1, nsteps + 2)

    vi, vj = numpy.meshgrid(i, j)
    vi.shape = -1
    vj.shape = -1

    ri = ((vi + sinV * vj) ** 2 + (cosV * vj ** 2))
    select = ri < (radius / spacing) ** 2

    xi = vi[select]
    yj = vj[select]

    offsets = xi[:, None] * d1 + yj[:, None] * d2
    offsets *= spacing
    gauss = numpy.exp(-(ri[select] * (spacing ** 2) / ((beam_waist) ** 2)))

    ray_data = numpy.zeros(offsets.shape[0], dtype=ray_dtype)
    P1 = (E1_amp * E1_amp.conjugate()).real
    P2 = (E2_amp * E1_amp.conjugate()).real
    total_power = ((P1 + P2) * (gauss ** 2)).sum()
    scaling = numpy.sqrt(beam_power / total_power)

    ray_data['origin'] = offsets
    ray_data['origin'] += origin
    ray_data['direction'] = direction
    ray_data['wavelength_idx'] = 0
    ray_data['E_vector'] = [normaliseVector(E_vector)]
    ray_data['E1_amp'] = scaling * E1_amp * gauss
    ray_data['E2_amp'] = scaling * E2_amp * gauss
    ray_data['refractive_index'] = 1.0 + 0.0j
    ray_data['normal'] = [[0, 1, 0]]
    ray_data['ray_type_id'] = (1 << 1)

    g = _from_rays(ray_data)
    wl = numpy.array([wavelength])
    return oracle.config_parabasal_rays(g, wl, spacing / blending, 0.0)


def make_hexagonal_grid(radius, spacing=1.0):
    cosV = numpy.cos(numpy.pi * 30 / 180.)
    sinV = numpy.sin(numpy.pi * 30 / 180.)
    nsteps = int(radius / (spacing * numpy.cos(numpy.pi * 30 / 180.)))
    i = numpy.arange(-nsteps - 1, nsteps + 2)
    j = numpy.arange(-nsteps - 1, nsteps + 2)
    vi, vj = numpy.meshgrid(i, j)
    vi.shape = -1
    vj.shape = -1
    x = (vi + sinV * vj) * spacing
    y = cosV * vj * spacing
    r2 = x ** 2 + y ** 2
    select = r2 < (radius ** 2)
    return x[select], y[select]


def gaussian_point_source(origin, direction, resolution, working_dist, numerical_aperture, blending, beam_waist,
                          wavelength, E_vector, E1_amp, E2_amp):
    origin = numpy.array(origin)
    direction = normaliseVector(numpy.array(direction))

    d2 = normaliseVector(numpy.cross(direction, E_vector))
    d1 = numpy.cross(direction, d2)

    focal_radius = beam_waist / 1000.0
    k_abs = (2000.0 * numpy.pi) / wavelength
    k_radius = 2. / focal_radius
    k_max = numerical_aperture * k_abs
    k_grid_spacing = k_max / resolution

    kx, ky = make_hexagonal_grid(k_max, spacing=k_grid_spacing)

    krsq = kx ** 2 + ky ** 2
    kz = numpy.sqrt(k_abs ** 2 - krsq)

    waist_location = origin + working_dist * direction
    directions = (kx[:, None] * d1 + ky[:, None] * d2 + kz[:, None] * direction) / k_abs
    directions = normaliseVector(directions)

    gauss = numpy.exp(-(krsq / (k_radius ** 2)))
    E1 = gauss * E1_amp
    E2 = gauss * E2_amp

    ray_data = numpy.zeros(gauss.shape[0], dtype=ray_dtype)
    E_vectors = numpy.cross(directions, d2[None, :])

    ray_data['origin'] = waist_location[None, :]
    ray_data['direction'] = directions
    ray_data['wavelength_idx'] = 0
    ray_data['E_vector'] = E_vectors
    ray_data['E1_amp'] = E1
    ray_data['E2_amp'] = E2
    ray_data['refractive_index'] = 1.0 + 0.0j
    ray_data['normal'] = [[0, 1, 0]]
    ray_data['ray_type_id'] = (1 << 1)

    g = _from_rays(ray_data)
    wl = numpy.array([wavelength])
    gauss_radius = 2. * blending / (k_grid_spacing)
    g = oracle.config_parabasal_rays(g, wl, gauss_radius, 0.0)
    return oracle.project_to_plane(g, origin, direction)
