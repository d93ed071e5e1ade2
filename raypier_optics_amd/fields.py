"""E-field evaluation from gausslets (drop-in for raypier.core.fields)."""
import numpy

from .cfields import sum_gaussian_modes, evaluate_modes as evaluate_modes_c, gausslet_modes
from .ctracer import RayCollection, GaussletCollection


def evaluate_neighbours_gc(gausslets):
    """(base rays, x, y, dx, dy) of a gausslet_dtype array: each parabasal
    ray's offset from its base origin and its slopes, in the frame
    (E, E x d, d) of the base ray (the quantities of
    core/fields.py:114-136).  The device path (rpx_gausslet_modes) computes
    the same values inside the fused mode fit; this host form serves callers
    that want them separately."""
    base = gausslets["base_ray"]
    e = base["E_vector"]
    d = base["direction"]
    frame = numpy.stack([e, numpy.cross(e, d), d], axis=1)          # (N, 3, 3), rows E, H, d
    offs = numpy.einsum("nij,npj->npi", frame, gausslets["para_rays"]["origin"] - base["origin"][:, None, :])
    dirs = numpy.einsum("nij,npj->npi", frame, gausslets["para_rays"]["direction"])
    return base, offs[..., 0], offs[..., 1], dirs[..., 0] / dirs[..., 2], dirs[..., 1] / dirs[..., 2]


def _base_collection(gc):
    """The base rays of a GaussletCollection as a RayCollection sharing its
    device buffer (the SoA ray rows are identical)."""
    dev = gc._ensure_device()
    rc = RayCollection.from_device(dev, gc.n_rays)
    rc.wavelengths = gc.wavelengths
    return rc


class EFieldSummation(object):
    """Mode coefficients computed once, field evaluated at many point sets
    (core/fields.py:221-249)."""

    def __init__(self, gausslet_collection, wavelengths=None, blending=1.0):
        if wavelengths is None:
            wavelengths = numpy.asarray(gausslet_collection.wavelengths)
        if wavelengths is None:
            raise ValueError("No wavelengths supplied")
        self.wavelengths = wavelengths
        self.gc = gausslet_collection
        self.modes = gausslet_modes(gausslet_collection, blending)
        self.base_rays = _base_collection(gausslet_collection)

    def evaluate(self, points, time_ps=0.0):
        points = numpy.ascontiguousarray(points)
        shape = points.shape
        E = sum_gaussian_modes(self.base_rays, self.modes, self.wavelengths, points.reshape(-1, 3), time_ps)
        return E.reshape(shape)


def eval_Efield_from_gausslets(gausslet_collection, points, wavelengths=None, blending=1.0, time_ps=0.0,
                               **kwds):
    """Vector E-field at `points` from a GaussletCollection
    (core/fields.py:252-276); complex128 (N, 3)."""
    if wavelengths is None:
        wavelengths = numpy.asarray(gausslet_collection.wavelengths)
    modes = gausslet_modes(gausslet_collection, blending)
    rays = _base_collection(gausslet_collection)
    return sum_gaussian_modes(rays, modes, wavelengths, points, time_ps, precision=kwds.get("precision", 0))


# ------------------------------------------------------------ ray fields
# eval_Efield_from_rays (core/fields.py:196-218) on the GPU: the optional
# projection to the exit-pupil sphere, the neighbour projection and mode fit
# of every fully surrounded ray, and the field sum all run on the rays where
# they lie in HBM.
import ctypes

from . import _abi
from ._lib import load, check, require_gpu, stream_handle


def _project_device(rc_dev, n, centre, radius):
    from ._device import DeviceRays
    torch = require_gpu()
    lib = load()
    out = DeviceRays(max(n, 1))
    ws = torch.empty(max(int(lib.rpx_trace_workspace_bytes(n)), 8), dtype=torch.uint8, device="cuda")
    k = ctypes.c_uint64(0)
    c = (ctypes.c_double * 3)(*[float(v) for v in numpy.asarray(centre, dtype=numpy.float64).reshape(3)])
    s_in, s_out = rc_dev.struct(), out.struct()
    check(lib.rpx_project_to_sphere(ctypes.byref(s_in), n, c, float(radius), ctypes.byref(s_out), ctypes.byref(k),
                                    ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream_handle()))
    return out, k.value


def _ray_modes_device(rc_dev, n, nb, blending, want_xy=False):
    from ._device import DeviceRays
    torch = require_gpu()
    lib = load()
    nb = nb.to(device="cuda", dtype=torch.int32).contiguous() if hasattr(nb, "data_ptr") else \
        torch.from_numpy(numpy.ascontiguousarray(nb, dtype=numpy.int32)).to("cuda")
    if nb.dim() != 2 or nb.shape[1] != 6 or nb.shape[0] != n:
        raise IndexError("neighbour table of shape %s for %d rays" % (tuple(nb.shape), n))
    out = DeviceRays(max(n, 1))
    modes = torch.empty((max(n, 1), 6), dtype=torch.float64, device="cuda")
    xy = torch.empty((max(n, 1), 24), dtype=torch.float64, device="cuda") if want_xy else None
    ws = torch.empty(max(int(lib.rpx_trace_workspace_bytes(n)), 8), dtype=torch.uint8, device="cuda")
    k = ctypes.c_uint64(0)
    s_in, s_out = rc_dev.struct(), out.struct()
    check(lib.rpx_ray_field_modes(ctypes.byref(s_in), n, ctypes.c_void_p(nb.data_ptr()), float(blending),
                                  ctypes.byref(s_out), ctypes.c_void_p(modes.data_ptr()),
                                  ctypes.c_void_p(xy.data_ptr()) if xy is not None else None, ctypes.byref(k),
                                  ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream_handle()))
    return out, k.value, modes[: k.value], (xy[: k.value] if xy is not None else None)


def project_to_sphere(rays, centre=(0, 0, 0), radius=10.0):
    """Rays (ray_dtype array) moved back to their intercept with a sphere;
    rays that miss it are dropped (core/fields.py:51-80)."""
    rc = RayCollection.from_array(numpy.ascontiguousarray(rays))
    if rc.n_rays == 0:
        return rc.copy_as_array()
    out, k = _project_device(rc._ensure_device(), rc.n_rays, centre, radius)
    return out.download(k)


def evaluate_neighbours(rays, neighbours_idx):
    """(rays with six neighbours, x, y, dx, dy) -- neighbour offsets on each
    ray's origin plane and slopes in its (E, E x d) frame, (N, 6) each
    (core/fields.py:83-111)."""
    rc = RayCollection.from_array(numpy.ascontiguousarray(rays))
    n = rc.n_rays
    out, k, _, xy = _ray_modes_device(rc._ensure_device(), n, neighbours_idx, 1.0, want_xy=True)
    xy = xy.cpu().numpy()
    return (out.download(k), xy[:, 0:6].copy(), xy[:, 6:12].copy(), xy[:, 12:18].copy(), xy[:, 18:24].copy())


def eval_Efield_from_rays(ray_collection, points, wavelengths, blending=1.0, time_ps=0.0, exit_pupil_offset=0.0,
                          exit_pupil_centre=(0.0, 0.0, 0.0)):
    """Vector E-field at `points` (N, 3) from a ray field with neighbours
    (core/fields.py:196-218); complex128 (N, 3)."""
    n = ray_collection.n_rays
    dev = ray_collection._ensure_device()
    nb = ray_collection._neighbours_device() if hasattr(ray_collection, "_neighbours_device") else None
    if nb is None:
        raise TypeError("the ray collection has no neighbour table")
    radius = exit_pupil_offset
    if radius:
        dev, n_p = _project_device(dev, n, exit_pupil_centre, radius)
        if n_p != n:
            # the neighbour table indexes the unprojected rays; the reference
            # fails on the length mismatch as well
            raise IndexError("%d of %d rays miss the exit-pupil sphere" % (n - n_p, n))
    out, k, modes, _ = _ray_modes_device(dev, n, nb, blending)
    rays = RayCollection.from_device(out, k)
    return sum_gaussian_modes(rays, modes, wavelengths, points, time_ps)
