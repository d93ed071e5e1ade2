"""Ray-source helpers that work on traced generations.

`set_face_sequence` is BaseRaySource.set_face_sequence (sources.py:138-155):
the face sequence for sequential tracing (tracer.trace_ray_sequence), taken
as the most common end face of each traced generation.  The per-generation
end-face histogram is one GPU launch over all generations
(rpx_face_histogram); the vote itself is read off the histogram.
"""
import numpy as np

from . import _abi
from ._lib import load, check, require_gpu, stream_handle
from .results import generation_table

NO_FACE = 0xFFFFFFFF


def end_face_histogram(traced_rays, n_bins):
    """uint64 (n_generations, n_bins + 2): counts of end_face_idx == b for
    b < n_bins, then of 0xFFFFFFFF, then of every other value."""
    torch = require_gpu()
    lib = load()
    gens, devs = generation_table(traced_rays)
    ws = torch.empty(int(lib.rpx_lineage_workspace_bytes(len(traced_rays), 0)), dtype=torch.uint8, device="cuda")
    counts = torch.empty((len(traced_rays), n_bins + 2), dtype=torch.int64, device="cuda")
    check(lib.rpx_face_histogram(gens, len(traced_rays), n_bins, counts.data_ptr(), ws.data_ptr(), ws.numel(),
                                 stream_handle()))
    return counts.cpu().numpy().astype(np.uint64)


def _most_common(counts, n_bins, rays):
    """The value numpy.unique(end_face_idx, return_counts=True) puts first
    among those with the largest count (the smallest such value)."""
    if counts.sum() == 0:
        raise ValueError("attempt to get argmax of an empty sequence")
    if counts[n_bins + 1]:
        # values other than 0..n_bins-1 and 0xFFFFFFFF: count them exactly
        ids, c = np.unique(rays.base_rays.end_face_idx, return_counts=True)
        return int(ids[c.argmax()])
    k = int(np.argmax(counts[: n_bins + 1]))
    return k if k < n_bins else NO_FACE


def set_face_sequence(traced_rays, face_lists, all_faces):
    """[(FaceList, face index in it), ...] for the generations of
    traced_rays, stopping at the first generation whose most common end face
    is not a face of face_lists."""
    global_map = {face: i for i, face in enumerate(all_faces)}
    face_map = {global_map[face]: (fl, face_idx) for fl in face_lists for face_idx, face in enumerate(fl.faces)}
    if not traced_rays:
        return []
    n_bins = len(all_faces)
    if n_bins > 4094:
        raise ValueError("more than 4094 faces")
    counts = end_face_histogram(traced_rays, n_bins)
    face_sequence = []
    for g, rays in enumerate(traced_rays):
        most_common = _most_common(counts[g], n_bins, rays)
        try:
            face_tuple = face_map[most_common]
        except KeyError:
            break
        face_sequence.append(face_tuple)
    return face_sequence


# ---------------------------------------------------------------- sources
# The lattice sources of sources.py / gausslet_sources.py, generated in HBM.
# Scalars and vectors are set up on the host exactly as the reference's
# input_rays properties compute them (numpy); every per-ray expression runs
# in the rpx_lattice_* kernels.
import ctypes
import math
import time


def _normalise(a):
    """normaliseVector (core/utils.py:5-9)."""
    a = np.asarray(a)
    mag = np.sqrt((a ** 2).sum(axis=-1))[..., np.newaxis]
    return a / mag


def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    return tuple(float(x) for x in v / math.sqrt(float((v * v).sum())))


def _lattice(kind, nsteps, limit, spacing, origin, d1, d2, direction, E_vector, weight_num, weight_den, E1, E2,
             ray_type_id, ray_ident_index, k_abs=0.0, k_abs_sq=0.0, power=0.0):
    L = _abi.Lattice()
    L.kind = kind
    L.nsteps = int(nsteps)
    L.cos30 = float(np.cos(np.pi * 30 / 180.))
    L.sin30 = float(np.sin(np.pi * 30 / 180.))
    L.limit = float(limit)
    L.spacing = float(spacing)
    for name, v in (("origin", origin), ("d1", d1), ("d2", d2), ("direction", direction), ("E_vector", E_vector)):
        arr = getattr(L, name)
        for k in range(3):
            arr[k] = float(v[k])
    L.weight_num, L.weight_den = float(weight_num), float(weight_den)
    L.E1[0], L.E1[1] = float(complex(E1).real), float(complex(E1).imag)
    L.E2[0], L.E2[1] = float(complex(E2).real), float(complex(E2).imag)
    L.k_abs, L.k_abs_sq, L.power = float(k_abs), float(k_abs_sq), float(power)
    L.ray_type_id = int(ray_type_id)
    L.ray_ident_index = int(ray_ident_index)
    return L


def _select(L):
    torch = require_gpu()
    lib = load()
    ws = torch.empty(int(lib.rpx_lattice_workspace_bytes(L.nsteps)), dtype=torch.uint8, device="cuda")
    n = ctypes.c_uint64(0)
    power = ctypes.c_double(0.0)
    check(lib.rpx_lattice_select(ctypes.byref(L), ctypes.byref(n), ctypes.byref(power),
                                 ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream_handle()))
    return n.value, power.value, ws


def _emit(L, n, ws, gausslet, neighbours=False):
    from ._device import DeviceRays
    torch = require_gpu()
    lib = load()
    dev = DeviceRays(max(n, 1), gausslet=gausslet)
    s = dev.struct()
    nb = torch.empty((max(n, 1), 6), dtype=torch.int32, device="cuda") if neighbours else None
    check(lib.rpx_lattice_emit(ctypes.byref(L), ctypes.byref(s), ctypes.c_void_p(nb.data_ptr()) if nb is not None
                               else None, ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream_handle()))
    return dev, (nb[:n] if nb is not None else None)


def _transverse_axes(direction):
    max_axis = np.abs(direction).argmax()
    if max_axis == 0:
        v = np.array([0., 1., 0.])
    else:
        v = np.array([1., 0., 0.])
    d1 = _normalise(np.cross(direction, v))
    d2 = _normalise(np.cross(direction, d1))
    return d1, d2


class _Source(object):
    _defaults = {}

    def __init__(self, **kw):
        for k, v in self._defaults.items():
            setattr(self, k, kw.pop(k, v))
        if kw:
            raise TypeError("unknown source parameters: %s" % ", ".join(sorted(kw)))
        self.direction = _unit(self.direction)
        self.E_vector = _unit(self.E_vector)

    @property
    def wavelength_list(self):
        return [self.wavelength]


class HexagonalRayFieldSource(_Source):
    """Rays on a hexagonal lattice of `resolution` steps across `radius`,
    Gaussian-weighted by gauss_width, with their six lattice neighbours
    (sources.py:1007-1107).  input_rays is a RayCollection in HBM."""
    _defaults = dict(origin=(0., 0., 0.), direction=(0., 0., 1.), radius=10., resolution=10.0, gauss_width=2.0,
                     E_vector=(1., 0., 0.), E1_amp=1.0 + 0.0j, E2_amp=0.0 + 0.0j, wavelength=0.78, max_ray_len=200.0)

    @property
    def input_rays(self):
        from .ctracer import RayCollection
        origin = np.array(self.origin)
        direction = np.array(self.direction)
        radius = self.radius
        spacing = radius / self.resolution
        d1, d2 = _transverse_axes(direction)
        E_vector = np.cross(self.E_vector, direction)
        E_vector = np.cross(E_vector, direction)
        cosV = np.cos(np.pi * 30 / 180.)
        sinV = np.sin(np.pi * 30 / 180.)
        d2 = cosV * d2 + sinV * d1
        nsteps = int(radius / (spacing * np.cos(np.pi * 30 / 180.)))
        L = _lattice(_abi.LATTICE_RAY_FIELD, nsteps, (radius / spacing) ** 2, spacing, origin, d1, d2, direction,
                     _normalise(E_vector), spacing ** 2, self.gauss_width ** 2, self.E1_amp, self.E2_amp, 0, 1)
        n, _, ws = _select(L)
        dev, nb = _emit(L, n, ws, False, neighbours=True)
        rc = RayCollection.from_device(dev, n)
        rc.wavelengths = self.wavelength_list
        rc._nb_dev = nb
        rc._mtime = time.monotonic()
        return rc

    @property
    def neighbours(self):
        return self.input_rays.neighbours


class CollimatedGaussletSource(_Source):
    """A quasi-collimated beam of gausslets on a hexagonal lattice, scaled to
    beam_power, parabasal rays of radius spacing / blending
    (gausslet_sources.py:357-457).  input_rays is a GaussletCollection in HBM."""
    _defaults = dict(origin=(0., 0., 0.), direction=(0., 0., 1.), radius=10., resolution=10.0, blending=1.5,
                     beam_waist=1.0, E_vector=(1., 0., 0.), E1_amp=1.0 + 0.0j, E2_amp=0.0 + 0.0j, wavelength=0.78,
                     beam_power=1.0, max_ray_len=200.0)

    @property
    def input_rays(self):
        from .ctracer import GaussletCollection
        origin = np.array(self.origin)
        direction = np.array(self.direction)
        radius = self.radius
        spacing = radius / self.resolution
        d1, d2 = _transverse_axes(direction)
        E_vector = np.cross(self.E_vector, direction)
        E_vector = np.cross(E_vector, direction)
        cosV = np.cos(np.pi * 30 / 180.)
        sinV = np.sin(np.pi * 30 / 180.)
        d2 = cosV * d2 + sinV * d1
        nsteps = int(radius / (spacing * np.cos(np.pi * 30 / 180.)))
        E1_amp, E2_amp = complex(self.E1_amp), complex(self.E2_amp)
        P1 = (E1_amp * E1_amp.conjugate()).real
        P2 = (E2_amp * E1_amp.conjugate()).real
        L = _lattice(_abi.LATTICE_RAY_FIELD, nsteps, (radius / spacing) ** 2, spacing, origin, d1, d2, direction,
                     _normalise(E_vector), spacing ** 2, self.beam_waist ** 2, 0j, 0j, 1 << 1, 0, power=P1 + P2)
        n, total_power, ws = _select(L)
        scaling = np.sqrt(self.beam_power / total_power)
        E1c, E2c = scaling * E1_amp, scaling * E2_amp
        L.E1[0], L.E1[1] = float(E1c.real), float(E1c.imag)
        L.E2[0], L.E2[1] = float(E2c.real), float(E2c.imag)
        dev, _ = _emit(L, n, ws, True)
        gc = GaussletCollection.from_device(dev, n)
        wl = np.array(self.wavelength_list)
        gc.wavelengths = wl
        gc.config_parabasal_rays(wl, spacing / self.blending, 0.0)
        return gc


class GaussianPointSource(_Source):
    """A low-F# point source of gausslets with a Gaussian angular profile,
    focused working_dist in front of the origin (gausslet_sources.py:460-566).
    input_rays is a GaussletCollection in HBM."""
    _defaults = dict(origin=(0., 0., 0.), direction=(0., 0., 1.), resolution=10.0, working_dist=100.0,
                     numerical_aperture=0.2, blending=1.0, beam_waist=1.0, E_vector=(1., 0., 0.),
                     E1_amp=1.0 + 0.0j, E2_amp=0.0 + 0.0j, wavelength=0.78, max_ray_len=200.0)

    @property
    def input_rays(self):
        from .ctracer import GaussletCollection
        origin = np.array(self.origin)
        direction = _normalise(np.array(self.direction))
        d2 = _normalise(np.cross(direction, self.E_vector))
        d1 = np.cross(direction, d2)
        focal_radius = self.beam_waist / 1000.0
        k_abs = (2000.0 * np.pi) / self.wavelength
        k_radius = 2. / focal_radius
        k_max = self.numerical_aperture * k_abs
        k_grid_spacing = k_max / self.resolution
        # make_hexagonal_grid(k_max, spacing=k_grid_spacing) (core/gausslets.py:55-80)
        nsteps = int(k_max / (k_grid_spacing * np.cos(np.pi * 30 / 180.)))
        waist_location = origin + self.working_dist * direction
        L = _lattice(_abi.LATTICE_POINT_SOURCE, nsteps, k_max ** 2, k_grid_spacing, waist_location, d1, d2,
                     direction, (0.0, 0.0, 0.0), 1.0, k_radius ** 2, self.E1_amp, self.E2_amp, 1 << 1, 0,
                     k_abs=k_abs, k_abs_sq=k_abs ** 2)
        n, _, ws = _select(L)
        dev, _ = _emit(L, n, ws, True)
        gc = GaussletCollection.from_device(dev, n)
        wl = np.array(self.wavelength_list)
        gc.wavelengths = wl
        gauss_radius = 2. * self.blending / (k_grid_spacing)
        gc.config_parabasal_rays(wl, gauss_radius, 0.0)
        gc.project_to_plane(origin, direction)
        return gc
