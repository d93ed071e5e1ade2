"""Beam decomposition (drop-in for raypier.core.gausslets).

decompose_position (core/gausslets.py:219-357) re-launches an arbitrary
gausslet beam as a hexagonal grid of new gausslets on a plane:

  1. the beam's field on a square grid of the plane -- EFieldSummation, on
     the GPU (rpx_gausslet_modes + rpx_sum_gaussian_modes);
  2. the phase of the dominant polarisation, optionally with a spherical
     reference removed, unwrapped (unwrap2d) and fitted by a smoothing
     bicubic spline (scipy) -- small host arrays, as in the reference;
  3. at the hexagonal launch points: slopes and curvatures of the spline,
     the mode bases and curvature coefficients (rpx_calc_mode_curvature),
     the field again (GPU), the new gausslets' parabasal rays and curvature
     (rpx_config_parabasal_rays, rpx_apply_mode_curvature), and a power
     normalisation against their own summed field (GPU).

decompose_angle (core/gausslets.py:97-216) launches gausslets over a grid of
directions from the FFT of a sampled field (host FFT and spline, device
parabasal set-up).
"""
import numpy as np

from .ctracer import GaussletCollection
from ._abi import ray_dtype
from .fields import EFieldSummation, eval_Efield_from_gausslets
from .cfields import calc_mode_curvature, apply_mode_curvature, build_interaction_matrix
from .unwrap2d import unwrap2d

root_pi = np.sqrt(np.pi)


def _normalise(a):
    a = np.asarray(a)
    return a / np.sqrt((a ** 2).sum(axis=-1))[..., np.newaxis]


def next_power_of_two(n):
    return int(2 ** np.ceil(np.log2(n)))


def lagrange_invariant(gausslet):
    """The six antisymmetric products h_i.u_j - h_j.u_i of a gausslet's
    parabasal offsets and directions in its (E, E x d) frame
    (core/gausslets.py:20-52)."""
    b = gausslet.base_ray
    origin = np.asarray(b.origin)
    d = np.asarray(b.direction)
    e1 = np.asarray(b.E_vector)
    e2 = np.cross(e1, d)
    paras = gausslet.parabasal_rays
    h = [((np.asarray(p.origin) - origin) * e1).sum() for p in paras], \
        [((np.asarray(p.origin) - origin) * e2).sum() for p in paras]
    u = [(np.asarray(p.direction) * e1).sum() for p in paras], [(np.asarray(p.direction) * e2).sum() for p in paras]

    def hu(i, j):
        return (h[0][i] * u[0][j] + h[1][i] * u[1][j]) - (h[0][j] * u[0][i] + h[1][j] * u[1][i])
    return [hu(0, 5), hu(1, 2), hu(3, 4), hu(1, 4), hu(0, 3), hu(2, 5)]


def make_hexagonal_grid(radius, spacing=1.0, connectivity=False):
    """Points of a hexagonal lattice inside `radius` (core/gausslets.py:55-94);
    with connectivity, also their six neighbours (-1 where unselected)."""
    cos30 = np.cos(np.pi * 30 / 180.)
    sin30 = np.sin(np.pi * 30 / 180.)
    nsteps = int(radius / (spacing * np.cos(np.pi * 30 / 180.)))
    idx = np.arange(-nsteps - 1, nsteps + 2)
    vi, vj = np.meshgrid(idx, idx)
    ni, nj = vi.shape
    vi, vj = vi.reshape(-1), vj.reshape(-1)
    x = (vi + sin30 * vj) * spacing
    y = cos30 * vj * spacing
    select = (x ** 2 + y ** 2) < (radius ** 2)
    if not connectivity:
        return x[select], y[select]
    labels = np.full((ni + 2, nj + 2), -1, np.int64)
    labels[1:-1, 1:-1] = np.arange(ni * nj).reshape(ni, nj)
    nb = np.dstack([labels[2:, 1:-1], labels[:-2, 1:-1], labels[1:-1, 2:], labels[1:-1, :-2],
                    labels[2:, :-2], labels[:-2, 2:]]).reshape(-1, 6)
    centres = np.full(vi.size, -1, np.int64)
    centres[select] = np.arange(select.sum())
    # the -1 entries index the last lattice point's slot, as in the reference
    return x[select], y[select], centres[nb[select]]


def decompose_position(input_rays, origin, direction, axis1, radius, resolution, curvature=None, blending=1.5):
    """(gausslets, E_field on the sampling grid, unwrapped phase)."""
    from scipy.interpolate import RectBivariateSpline
    spacing = radius / resolution
    wavelengths = input_rays.wavelengths
    if len(np.unique(wavelengths)) > 1:
        raise ValueError("Can't decompose wavefront with more than one wavelength present.")
    wavelength = wavelengths[0]
    origin = np.asarray(origin)
    direction = _normalise(np.asarray(direction))
    axis1 = _normalise(np.asarray(axis1))
    axis2 = _normalise(np.cross(axis1, direction))
    axis1 = np.cross(axis2, direction)

    # field on a square grid a little larger than the launch disc
    _radius = radius * (1 + 2. / resolution)
    grid = np.linspace(-_radius, _radius, int((1.1 * resolution) * 2))
    x, y = np.meshgrid(grid, grid)
    samples = origin[None, :] + x.reshape(-1, )[:, None] * axis1[None, :] + y.reshape(-1)[:, None] * axis2[None, :]
    efield = EFieldSummation(input_rays)
    E_field = efield.evaluate(samples)

    # phase of the stronger polarisation
    E1 = (E_field * axis1[None, :]).sum(axis=1)
    E2 = (E_field * axis2[None, :]).sum(axis=1)
    P1 = (E1.real ** 2 + E1.imag ** 2).sum()
    P2 = (E2.real ** 2 + E2.imag ** 2).sum()
    main = E1 if P1 > P2 else E2
    phase = np.arctan2(main.imag, main.real)
    nx, ny = len(grid), len(grid)
    phase.shape = (nx, ny)
    if curvature is not None:
        sphz = -(curvature - np.sqrt(curvature * curvature - x * x - y * y))
        sph = sphz * 2000.0 * np.pi / wavelength - np.pi
        phase = phase - sph
        phase = phase % (2 * np.pi)
        phase = phase - np.pi
    else:
        sph = 0
    uphase, residuals = unwrap2d(phase, anchor=(nx // 2, ny // 2))
    uphase2 = uphase + sph
    wavefront = RectBivariateSpline(grid, grid, -uphase2 * wavelength / (2000 * np.pi), kx=3, ky=3, s=0.001)

    # the new gausslets on a hexagonal grid
    rx, ry, nb = make_hexagonal_grid(radius, spacing=spacing, connectivity=True)
    N = len(rx)
    origins = origin[None, :] + rx[:, None] * axis1[None, :] + ry[:, None] * axis2[None, :]
    dx = wavefront(rx, ry, dx=1, grid=False)
    dy = wavefront(rx, ry, dy=1, grid=False)
    dx2 = wavefront(rx, ry, dx=2, grid=False)
    dy2 = wavefront(rx, ry, dy=2, grid=False)
    dxdy = wavefront(rx, ry, dx=1, dy=1, grid=False)
    A, B, C, xl, yl, zl = calc_mode_curvature(rx, ry, dx, dy, dx2, dy2, dxdy)
    E_in = efield.evaluate(origins)
    directions = axis1[None, :] * zl[:, 0, None] + axis2[None, :] * zl[:, 1, None] + direction[None, :] * zl[:, 2, None]
    E_vectors = axis1[None, :] * xl[:, 0, None] + axis2[None, :] * xl[:, 1, None] + direction[None, :] * xl[:, 2, None]
    H_vectors = axis1[None, :] * yl[:, 0, None] + axis2[None, :] * yl[:, 1, None] + direction[None, :] * yl[:, 2, None]

    ray_data = np.zeros(N, dtype=ray_dtype)
    ray_data['origin'] = origins
    ray_data['direction'] = directions
    ray_data['E_vector'] = E_vectors
    ray_data['refractive_index'] = 1.0
    ray_data['E1_amp'] = (E_in * E_vectors).sum(axis=-1)
    ray_data['E2_amp'] = (E_in * H_vectors).sum(axis=-1)

    gausslets = GaussletCollection.from_rays(ray_data)
    gausslets.wavelengths = wavelengths
    gausslets.config_parabasal_rays(wavelengths, spacing / blending, 0.0)
    apply_mode_curvature(gausslets, -A, -B, -C)

    E_test = eval_Efield_from_gausslets(gausslets, origins)
    power_scaling = (E_in.real ** 2 + E_in.imag ** 2).sum() / (E_test.real ** 2 + E_test.imag ** 2).sum()
    gausslets.scale_amplitude(np.sqrt(power_scaling))
    return gausslets, E_field, uphase


def decompose_angle(origin, direction, axis1, E_field, input_spacing, max_angle, wavelength, oversample=4,
                    E_max=None, pos_max=None):
    """(gausslets, angular spectrum): gausslets from `origin` over the
    directions within max_angle, amplitudes from the FFT of the sampled
    field E_field (N, M, 3) with spacing input_spacing (microns)."""
    from numpy import fft
    from scipy.interpolate import RectBivariateSpline
    input_spacing /= 1000.0
    wavelength /= 1000.0
    direction = _normalise(direction)
    d2 = _normalise(np.cross(direction, axis1))
    d1 = np.cross(direction, d2)
    N, M = E_field.shape[:2]
    size = next_power_of_two(max(N, M)) * oversample
    n0 = int((size / 2) - (N / 2))
    m0 = int((size / 2) - (M / 2))
    data_in = np.zeros((size, size, 3), dtype=np.complex128)
    data_in[n0:n0 + N, m0:m0 + M, :] = E_field
    data_in = fft.ifftshift(data_in, axes=(0, 1))
    data_out = fft.fft2(data_in, axes=(0, 1)) / (size ** 2)
    data_out = fft.fftshift(data_out, axes=(0, 1))

    kmax = 2.0 * np.pi / (2 * input_spacing)
    k_abs = 2.0 * np.pi / wavelength
    kr = np.linspace(-kmax, kmax, data_out.shape[0] + 1)[:-1]
    k_limit = k_abs * np.sin(np.pi * max_angle / 180.0)
    k_grid_spacing = (kr[1] - kr[0]) * oversample
    kx, ky = make_hexagonal_grid(k_limit, spacing=k_grid_spacing)
    kz = np.sqrt(k_abs ** 2 - kx ** 2 - ky ** 2)
    directions = (kx[:, None] * d1 + ky[:, None] * d2 + kz[:, None] * direction) / k_abs

    def spline(component):
        return RectBivariateSpline(kr, kr, component)(kx, ky, grid=False)
    E_real = np.column_stack([spline(data_out[:, :, q].real) for q in range(3)])
    E_imag = np.column_stack([spline(data_out[:, :, q].imag) for q in range(3)])
    E_full = -(-1.0j * E_real + E_imag)
    E_vectors = _normalise(np.cross(directions, d2))
    E2_vectors = _normalise(np.cross(directions, E_vectors))
    E1_amp = (E_vectors * E_full).sum(axis=1)
    E2_amp = (E2_vectors * E_full).sum(axis=1)

    gausslet_radius = wavelength * k_abs / (k_grid_spacing * np.pi)
    scaling = np.sqrt(2) * input_spacing * size
    ray_data = np.zeros(directions.shape[0], dtype=ray_dtype)
    ray_data['origin'] = origin + ((d1 + d2) * (input_spacing * 0.5))
    ray_data['direction'] = directions
    ray_data['wavelength_idx'] = 0
    ray_data['E_vector'] = E_vectors
    ray_data['E1_amp'] = E1_amp * scaling
    ray_data['E2_amp'] = E2_amp * scaling
    ray_data['refractive_index'] = 1.0 + 0.0j
    ray_data['normal'] = [[0, 1, 0]]
    ray_data['phase'] = 0.0
    wl = np.array([wavelength])
    rays = GaussletCollection.from_rays(ray_data)
    rays.wavelengths = wl
    rays.config_parabasal_rays(wl, gausslet_radius, 0.0)
    if E_max is not None and pos_max is not None:
        # the reference evaluates a trial field here but leaves the
        # amplitudes unscaled (its scaling lines are commented out)
        eval_Efield_from_gausslets(rays, pos_max[None, :], blending=1.0)
        rays = GaussletCollection.from_rays(ray_data)
        rays.wavelengths = wl
        rays.config_parabasal_rays(wl, gausslet_radius, 0.0)
    return rays, data_out
