"""Beam decomposition (drop-in for raypier.core.gausslets).

decompose_position(input_rays, origin, direction, axis1, radius, resolution,
curvature=None, blending=1.5) re-launches a gausslet beam as new gausslets
on a hexagonal lattice of the plane through `origin` normal to `direction`
(the role of core/gausslets.py:219-357); decompose_angle(...) launches
gausslets from one point over a lattice of directions taken from a sampled
field's angular spectrum (the role of core/gausslets.py:97-216).  Both are
designed here from the optics, not from the reference's steps:

decompose_position
  1. The beam's field on a square grid at half the launch spacing, a few
     samples beyond the launch disc -- EFieldSummation, on the GPU.
  2. The phase of the stronger transverse polarisation, with an optional
     reference sphere removed, unwrapped by least squares (unwrap2d) and
     turned into a wavefront W (mm of optical path, the sign the mode
     curvature kernels take).
  3. At each launch point a weighted quadratic least-squares fit of W over
     the surrounding 7 x 7 samples gives the slopes and curvatures that
     rpx_calc_mode_curvature turns into the launch direction, the local
     basis and the mode coefficients.
  4. Amplitudes are the field at the launch points (GPU) on that basis; the
     parabasal rays come from config_parabasal_rays and
     apply_mode_curvature (GPU); one complex scale, the least-squares fit of
     the new beam's field to the sampled field over the launch disc (GPU),
     fixes power and absolute phase together.

decompose_angle
  1. Every gausslet launched from `origin` has the waist w = 2 / dk (its
     far-field width is one lattice step dk), so in the sample plane the
     gausslets together are exp(-r^2 / w^2) times a sum of plane waves.
     The samples are therefore divided by that envelope, and the spectrum
     of the quotient is evaluated exactly at the hexagonal lattice of
     transverse wave vectors (spacing 2 pi / aperture, inside
     k sin(max_angle)) by a direct, separable Fourier sum -- no padding,
     FFT or interpolation.
  2. Each lattice vector becomes a gausslet with its plane-wave amplitude
     and parabasal rays from config_parabasal_rays (GPU).
  3. As above, one complex least-squares scale against the sampled field.
"""
import numpy as np

from .ctracer import GaussletCollection
from ._abi import ray_dtype
from .fields import EFieldSummation, eval_Efield_from_gausslets
from .cfields import calc_mode_curvature, apply_mode_curvature, build_interaction_matrix
from .unwrap2d import unwrap2d, wrap

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



def _plane_frame(direction, axis1):
    """Right-handed orthonormal (u, v, w): w along direction, u the part of
    axis1 normal to it, v = w x u."""
    w = _normalise(np.asarray(direction, np.float64))
    a = np.asarray(axis1, np.float64)
    u = _normalise(a - np.dot(a, w) * w)
    return u, np.cross(w, u), w


def _best_scale(e_model, e_target):
    """The complex s minimising sum |s e_model - e_target|^2."""
    num = np.vdot(e_model.reshape(-1), e_target.reshape(-1))
    den = np.vdot(e_model.reshape(-1), e_model.reshape(-1)).real
    return num / den if den > 0 else 1.0


def _local_quadratic(grid, W, px, py, half=3):
    """Slopes and second derivatives (Wx, Wy, Wxx, Wyy, Wxy) at (px, py) of
    W sampled on grid x grid (W[i, j] at x = grid[j], y = grid[i]): a
    Gaussian-weighted least-squares quadratic over the (2 half + 1)^2
    samples around the nearest grid node, weights of width 1.5 steps."""
    h = grid[1] - grid[0]
    n = len(grid)
    jc = np.clip(np.rint((px - grid[0]) / h).astype(np.int64), half, n - 1 - half)
    ic = np.clip(np.rint((py - grid[0]) / h).astype(np.int64), half, n - 1 - half)
    off = np.arange(-half, half + 1)
    oi, oj = np.meshgrid(off, off, indexing="ij")
    oi, oj = oi.reshape(-1), oj.reshape(-1)
    I = ic[:, None] + oi[None, :]
    J = jc[:, None] + oj[None, :]
    dx = grid[J] - px[:, None]
    dy = grid[I] - py[:, None]
    wt = np.exp(-(dx * dx + dy * dy) / (2.0 * (1.5 * h) ** 2))
    basis = np.stack([np.ones_like(dx), dx, dy, dx * dx, dx * dy, dy * dy], axis=-1)   # (P, S, 6)
    lhs = np.einsum("ps,psa,psb->pab", wt, basis, basis)
    rhs = np.einsum("ps,psa,ps->pa", wt, basis, W[I, J])
    c = np.linalg.solve(lhs, rhs[..., None])[..., 0]
    return c[:, 1], c[:, 2], 2.0 * c[:, 3], 2.0 * c[:, 5], c[:, 4]


def decompose_position(input_rays, origin, direction, axis1, radius, resolution, curvature=None, blending=1.5):
    """(gausslets, field on the sampling grid (n*n, 3), unwrapped phase
    (n, n)).  `curvature` R > 0 removes, before unwrapping, the reference
    sphere of a wave converging to a point R beyond the plane (R < 0:
    diverging from a point |R| before it); it is added back afterwards."""
    wavelengths = np.asarray(input_rays.wavelengths)
    if len(np.unique(wavelengths)) > 1:
        raise ValueError("Can't decompose wavefront with more than one wavelength present.")
    wl = float(wavelengths[0])
    k_mm = 2000.0 * np.pi / wl
    origin = np.asarray(origin, np.float64)
    u, v, w = _plane_frame(direction, axis1)
    spacing = radius / resolution

    # 1. samples at half the launch spacing, 4 steps beyond the disc
    h = 0.5 * spacing
    half_n = int(np.ceil(radius / h)) + 4
    grid = h * np.arange(-half_n, half_n + 1)
    X, Y = np.meshgrid(grid, grid)
    pts = origin + X.reshape(-1, 1) * u + Y.reshape(-1, 1) * v
    field = EFieldSummation(input_rays)
    E = field.evaluate(pts)

    # 2. phase of the stronger transverse component, unwrapped about the centre
    Eu, Ev = E @ u, E @ v
    main = Eu if np.vdot(Eu, Eu).real >= np.vdot(Ev, Ev).real else Ev
    psi = np.angle(main).reshape(X.shape)
    r2 = X * X + Y * Y
    if curvature is not None:
        Rc = float(curvature)
        sphere = -k_mm * np.sign(Rc) * (abs(Rc) - np.sqrt(np.maximum(Rc * Rc - r2, 0.0)))
    else:
        sphere = np.zeros_like(psi)
    c = half_n
    uphase, _ = unwrap2d(wrap(psi - sphere), anchor=(c, c))
    uphase = uphase + sphere
    W = -uphase / k_mm

    # 3. launch lattice and the local wavefront fit
    rx, ry = make_hexagonal_grid(radius, spacing=spacing)
    dx, dy, dx2, dy2, dxdy = _local_quadratic(grid, W, rx, ry)
    A, B, C, xl, yl, zl = calc_mode_curvature(rx, ry, dx, dy, dx2, dy2, dxdy)
    frame = np.stack([u, v, w])                    # rows: the plane's axes
    launch = origin + rx[:, None] * u + ry[:, None] * v
    e_vec, h_vec, d_vec = xl @ frame, yl @ frame, zl @ frame

    # 4. amplitudes, parabasal rays, one complex scale
    E_in = field.evaluate(launch)
    rays = np.zeros(len(rx), dtype=ray_dtype)
    rays["origin"] = launch
    rays["direction"] = d_vec
    rays["E_vector"] = e_vec
    rays["refractive_index"] = 1.0
    rays["E1_amp"] = np.einsum("ij,ij->i", E_in, e_vec)
    rays["E2_amp"] = np.einsum("ij,ij->i", E_in, h_vec)
    out = GaussletCollection.from_rays(rays)
    out.wavelengths = wavelengths
    out.config_parabasal_rays(wavelengths, spacing / blending, 0.0)
    apply_mode_curvature(out, -A, -B, -C)
    inside = r2.reshape(-1) <= (radius - spacing) ** 2
    s = _best_scale(eval_Efield_from_gausslets(out, pts[inside]), E[inside])
    out.scale_amplitude(s)
    return out, E, uphase


def decompose_angle(origin, direction, axis1, E_field, input_spacing, max_angle, wavelength, oversample=4,
                    E_max=None, pos_max=None):
    """(gausslets, angular spectrum at their wave vectors (K, 3)).

    E_field (N, M, 3) samples a field on the plane through `origin` normal
    to `direction`: sample [i, j] lies at origin + (i - N/2) s a + (j - M/2) s b
    with s = input_spacing (microns), b = normalize(direction x axis1) and
    a = b x direction.  wavelength in microns, max_angle in degrees.
    `oversample` sets the lattice step to 2 pi / (aperture * oversample / 4),
    so the default 4 gives one step per 2 pi / aperture; E_max and pos_max
    are accepted for compatibility and have no effect (the amplitude scale
    comes from the least-squares fit to the samples)."""
    s = input_spacing / 1000.0
    wl_mm = wavelength / 1000.0
    k = 2.0 * np.pi / wl_mm
    dvec = _normalise(np.asarray(direction, np.float64))
    b = _normalise(np.cross(dvec, axis1))
    a = np.cross(b, dvec)
    E_field = np.asarray(E_field, np.complex128)
    N, M = E_field.shape[:2]
    xs = (np.arange(N) - N / 2) * s
    ys = (np.arange(M) - M / 2) * s
    dk = 2.0 * np.pi / (max(N, M) * s) * (4.0 / oversample)
    kx, ky = make_hexagonal_grid(k * np.sin(np.radians(max_angle)), spacing=dk)
    kz = np.sqrt(k * k - kx * kx - ky * ky)
    waist = 2.0 / dk
    # 1. spectrum at the lattice of the samples over the gausslet envelope:
    #    sum_ij E[i, j] exp(r_ij^2 / w^2) exp(-i (kx x_i + ky y_j)) s^2
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    flat = E_field * np.exp((X * X + Y * Y) / (waist * waist))[..., None]
    fx = np.exp(-1j * np.outer(kx, xs))
    fy = np.exp(-1j * np.outer(ky, ys))
    spec = np.einsum("ki,ijc,kj->kc", fx, flat, fy) * (s * s)
    # 2. one gausslet per wave vector
    dirs = (kx[:, None] * a + ky[:, None] * b + kz[:, None] * dvec) / k
    e_vec = _normalise(a[None, :] - (dirs @ a)[:, None] * dirs)
    h_vec = np.cross(dirs, e_vec)
    rays = np.zeros(len(kx), dtype=ray_dtype)
    rays["origin"] = np.asarray(origin, np.float64)
    rays["direction"] = dirs
    rays["E_vector"] = e_vec
    rays["E1_amp"] = np.einsum("kc,kc->k", spec, e_vec)
    rays["E2_amp"] = np.einsum("kc,kc->k", spec, h_vec)
    rays["refractive_index"] = 1.0
    rays["normal"] = dvec
    wl = np.array([wavelength])
    out = GaussletCollection.from_rays(rays)
    out.wavelengths = wl
    out.config_parabasal_rays(wl, waist, 0.0)
    # 3. one complex scale against the samples
    pts = np.asarray(origin, np.float64) + X.reshape(-1, 1) * a + Y.reshape(-1, 1) * b
    sc = _best_scale(eval_Efield_from_gausslets(out, pts), E_field.reshape(-1, 3))
    out.scale_amplitude(sc)
    return out, spec * sc
