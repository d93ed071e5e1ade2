"""Differentiable spot RMS of a coaxial sequential lens (hand-written adjoint).

`SequentialLens` describes surfaces z = z_s + c r^2/(1+sqrt(1-(1+k)c^2 r^2))
+ A4 r^4 + A6 r^6 + A8 r^8 + A10 r^10 with per-surface parameters
(c, k, A4, A6, A8, A10), the media on either side of each surface and a
detector plane (include/rpx.h, rpx_lens_t).  `spot_rms(params, lens, rays)`
is a torch.autograd.Function: its forward pass runs the forward trace kernel
on the GPU, its backward pass the adjoint kernel, which recomputes every
intersection instead of storing it, and returns d(spot RMS)/d(params).

`SequentialLens.face_lists(params)` builds the same system from drop-in
faces (AsphericFace surfaces, FullDielectricDispersiveMaterial interfaces,
an opaque CircularFace detector) so that the spot RMS can be evaluated by
the float64 trace (tracer.trace_rays) as well -- which is how the gradient
is checked by central finite differences.
"""
import ctypes

import numpy as np

from . import _abi
from ._lib import load, check, require_gpu, stream_handle

N_PARAM = 6
PARAM_NAMES = ("c", "k", "A4", "A6", "A8", "A10")


class SequentialLens(object):
    def __init__(self, z, params, apertures, before, after, media, detector_z):
        self.z = np.asarray(z, np.float64)
        self.params = np.asarray(params, np.float64).reshape(-1, N_PARAM)
        self.apertures = np.asarray(apertures, np.float64)
        self.before = [int(x) for x in before]
        self.after = [int(x) for x in after]
        self.media = list(media)          # dispersion curves (BaseDispersionCurve-like)
        self.detector_z = float(detector_z)
        if not 1 <= len(self.z) <= _abi.LENS_MAX_SURFACES:
            raise ValueError("1..%d surfaces" % _abi.LENS_MAX_SURFACES)

    @property
    def n_surfaces(self):
        return len(self.z)

    @classmethod
    def from_prescription(cls, presc, media, detector_z):
        return cls([p["z"] for p in presc],
                   [[p["c"], p["k"]] + list(p["A"]) for p in presc],
                   [p["aperture"] for p in presc], [p["before"] for p in presc], [p["after"] for p in presc],
                   media, detector_z)

    @classmethod
    def from_workload(cls, w):
        """The lens of a workloads.lens_job_c5 job."""
        return cls.from_prescription(w.info["prescription"], w.info["media"], w.info["detector_z_mm"])

    def index_table(self, wavelengths):
        """Real refractive index of every medium at every wavelength:
        float64 (n_wavelengths, n_media)."""
        wl = np.asarray(wavelengths, np.float64)
        return np.stack([np.real(m.evaluate_n(wl)) for m in self.media], axis=1).astype(np.float64)

    def struct(self, params=None):
        P = self.params if params is None else np.asarray(params, np.float64).reshape(-1, N_PARAM)
        L = _abi.Lens()
        L.n_surfaces = self.n_surfaces
        L.n_media = len(self.media)
        L.detector_z = self.detector_z
        for s in range(self.n_surfaces):
            L.z[s] = self.z[s]
            for q in range(N_PARAM):
                L.param[s][q] = P[s, q]
            L.aperture[s] = self.apertures[s]
            L.medium_before[s] = self.before[s]
            L.medium_after[s] = self.after[s]
        return L

    def face_lists(self, params=None):
        """The lens as drop-in faces: [lens FaceList, detector FaceList]."""
        from .ctracer import FaceList
        from .cfaces import AsphericFace, CircularFace
        from .cshapes import CircleShape
        from .cmaterials import FullDielectricDispersiveMaterial, OpaqueMaterial
        P = self.params if params is None else np.asarray(params, np.float64).reshape(-1, N_PARAM)
        faces = []
        for s in range(self.n_surfaces):
            c, k, A4, A6, A8, A10 = (float(x) for x in P[s])
            # inverted normals: the material's outside is the -z side (before)
            mat = FullDielectricDispersiveMaterial(dispersion_inside=self.media[self.after[s]],
                                                   dispersion_outside=self.media[self.before[s]])
            faces.append(AsphericFace(shape=CircleShape(radius=float(self.apertures[s])), material=mat,
                                      conic_const=k, z_height=float(self.z[s]), curvature=-1.0 / c,
                                      A4=A4, A6=A6, A8=A8, A10=A10, invert_normals=True, atol=1e-16))
        lens = FaceList()
        lens.faces = faces
        det = FaceList()
        det.faces = [CircularFace(diameter=1e4, z_plane=self.detector_z, material=OpaqueMaterial())]
        return [lens, det]


class _Device(object):
    """Device inputs of the lens kernels for one ray collection."""

    def __init__(self, lens, rays):
        import torch
        self.rays = rays
        self.dev = rays._ensure_device()
        self.table = torch.from_numpy(lens.index_table(rays.wavelengths).reshape(-1)).to("cuda")
        self.n_wl = len(rays.wavelengths)
        lib = load()
        need = int(lib.rpx_lens_workspace_bytes(rays.n_rays, lens.n_surfaces))
        self.ws = torch.empty(max(need, 8), dtype=torch.uint8, device="cuda")


def _moments(lens, d, params):
    import torch
    lib = load()
    L = lens.struct(params)
    s = d.dev.struct()
    m = torch.empty(4, dtype=torch.float64, device="cuda")
    check(lib.rpx_lens_spot(ctypes.byref(L), ctypes.byref(s), d.rays.n_rays, ctypes.c_void_p(d.table.data_ptr()),
                            d.n_wl, ctypes.c_void_p(m.data_ptr()), ctypes.c_void_p(d.ws.data_ptr()), d.ws.numel(),
                            stream_handle()))
    return m


def _sigma(m):
    import torch
    c = m[1:3] / m[0]
    return torch.sqrt(torch.clamp(m[3] / m[0] - (c * c).sum(), min=0.0))


def spot_moments(lens, rays, params=None):
    """(kept rays, sum x, sum y, sum x^2+y^2) of the detector points, float64."""
    d = _Device(lens, rays)
    return _moments(lens, d, lens.params if params is None else params).cpu().numpy()


def spot_rms_and_grad(lens, rays, params=None):
    """Spot RMS (mm) and d(spot RMS)/d(params) as a (n_surfaces, 6) array."""
    import torch
    P = lens.params if params is None else np.asarray(params, np.float64).reshape(-1, N_PARAM)
    d = _Device(lens, rays)
    m = _moments(lens, d, P)
    g = _grad(lens, d, P, m)
    return float(_sigma(m)), g.cpu().numpy().reshape(-1, N_PARAM)


def _grad(lens, d, P, m):
    import torch
    lib = load()
    L = lens.struct(P)
    s = d.dev.struct()
    g = torch.empty(lens.n_surfaces * N_PARAM, dtype=torch.float64, device="cuda")
    check(lib.rpx_lens_spot_grad(ctypes.byref(L), ctypes.byref(s), d.rays.n_rays,
                                 ctypes.c_void_p(d.table.data_ptr()), d.n_wl, ctypes.c_void_p(m.data_ptr()),
                                 ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(d.ws.data_ptr()), d.ws.numel(),
                                 stream_handle()))
    return g


def _function():
    import torch

    class SpotRMS(torch.autograd.Function):
        """spot RMS (0-d tensor) of `lens` for `rays` as a function of the
        (n_surfaces, 6) float64 parameter tensor."""

        @staticmethod
        def forward(ctx, params, lens, rays):
            require_gpu()
            P = params.detach().to("cpu", torch.float64).numpy()
            d = _Device(lens, rays)
            m = _moments(lens, d, P)
            ctx.lens, ctx.d, ctx.P, ctx.m = lens, d, P, m
            ctx.param_device, ctx.param_dtype = params.device, params.dtype
            return _sigma(m).to(params.device, params.dtype)

        @staticmethod
        def backward(ctx, gout):
            g = _grad(ctx.lens, ctx.d, ctx.P, ctx.m).view(-1, N_PARAM)
            g = g.to(ctx.param_device, ctx.param_dtype) * gout
            return g, None, None

    return SpotRMS


_SPOT_RMS = None


def spot_rms(params, lens, rays):
    """Differentiable spot RMS: params is a (n_surfaces, 6) float64 tensor."""
    global _SPOT_RMS
    if _SPOT_RMS is None:
        _SPOT_RMS = _function()
    return _SPOT_RMS.apply(params, lens, rays)
