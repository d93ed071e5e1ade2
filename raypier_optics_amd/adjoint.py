"""Differentiable spot RMS and wavefront RMS of a coaxial sequential lens
(hand-written adjoint).

`SequentialLens` describes surfaces z = z_s + c r^2/(1+sqrt(1-(1+k)c^2 r^2))
+ A4 r^4 + A6 r^6 + A8 r^8 + A10 r^10 with per-surface parameters
(c, k, A4, A6, A8, A10), optionally plus a Zernike distortion
(ZernikeDistortion over ANSI indices j <= 27, n <= 6, as DistortionFace adds
it to the sag), the media on either side of each surface and a detector
plane (include/rpx.h, rpx_lens_t).

`spot_rms(params, lens, rays)` and `lens_objective(params, zcoefs, lens,
rays, objective)` are torch.autograd.Functions: the forward pass runs the
forward trace kernel on the GPU, the backward pass the adjoint kernel, which
recomputes every intersection instead of storing it, and returns
d(objective)/d(params) and d(objective)/d(Zernike coefficients).  The
objectives are the detector spot RMS and the wavefront RMS (optical path to
the on-axis detector point, piston removed per wavelength).

`SequentialLens.face_lists(params, zcoefs)` builds the same system from
drop-in faces (AsphericFace surfaces, DistortionFace over them for the
distorted ones, FullDielectricDispersiveMaterial interfaces, an opaque
CircularFace detector) so that the objectives can be evaluated by the
float64 trace (tracer.trace_rays) as well -- which is how the gradients are
checked by central finite differences.
"""
import ctypes
import math

import numpy as np

from . import _abi
from ._lib import load, check, require_gpu, stream_handle

N_PARAM = 6
PARAM_NAMES = ("c", "k", "A4", "A6", "A8", "A10")
ZDEG = 6
OBJECTIVES = {"spot": _abi.LENS_SPOT_RMS, "wavefront": _abi.LENS_WAVEFRONT_RMS}


def monomials():
    """(a, b) of the monomials x^a y^b in the device table's order: degree
    by degree, a descending."""
    return [(a, d - a) for d in range(ZDEG + 1) for a in range(d, -1, -1)]


def _poly_mul(p, q):
    out = {}
    for (a1, b1), c1 in p.items():
        for (a2, b2), c2 in q.items():
            k = (a1 + a2, b1 + b2)
            out[k] = out.get(k, 0) + c1 * c2
    return out


def zernike_monomials(j):
    """Cartesian expansion of one ZernikeDistortion term in the unit-radius
    coordinates: Z_j(x, y) = N R_n^|m|(r) cos(m theta) for m >= 0 and
    N R_n^|m|(r) sin(|m| theta) for m < 0 (cdistortions.pyx:456-514),
    N = sqrt(n+1) (m = 0) or sqrt(2(n+1)); returns the coefficients of
    monomials() as a float64 array.  R_n^|m| r^-|m| is a polynomial in
    r^2 = x^2 + y^2 and r^|m| (cos, sin)(|m| theta) = (Re, Im)(x + iy)^|m|."""
    from .cdistortions import eval_nmk
    n, m, _ = eval_nmk(int(j))
    am = abs(m)
    if n > ZDEG:
        raise ValueError("Zernike term j=%d has n=%d > %d" % (j, n, ZDEG))
    # (x + iy)^|m| = sum_k C(|m|, k) x^(|m|-k) (iy)^k
    harm = {}
    for k in range(am + 1):
        ik = [1, 1j, -1, -1j][k % 4]
        val = math.comb(am, k) * ik
        part = val.real if m >= 0 else val.imag
        if part:
            harm[(am - k, k)] = harm.get((am - k, k), 0) + int(round(part))
    radial = {}
    for k in range((n - am) // 2 + 1):
        a = (-1) ** k * math.factorial(n - k) // (math.factorial(k) * math.factorial((n + am) // 2 - k)
                                                   * math.factorial((n - am) // 2 - k))
        p = (n - 2 * k - am) // 2            # r^(2p) = (x^2 + y^2)^p
        for i in range(p + 1):
            key = (2 * i, 2 * (p - i))
            radial[key] = radial.get(key, 0) + a * math.comb(p, i)
    poly = _poly_mul(radial, harm)
    N = math.sqrt(float(n + 1)) if m == 0 else math.sqrt(float(2 * (n + 1)))
    idx = {ab: i for i, ab in enumerate(monomials())}
    out = np.zeros(_abi.LENS_ZTERMS)
    for ab, c in poly.items():
        if c:
            out[idx[ab]] += N * c
    return out


class SequentialLens(object):
    def __init__(self, z, params, apertures, before, after, media, detector_z, zernike=None):
        """zernike: one entry per surface, None or (unit_radius, [(ANSI j,
        coefficient), ...]) -- a ZernikeDistortion added to that surface's
        sag."""
        self.z = np.asarray(z, np.float64)
        self.params = np.asarray(params, np.float64).reshape(-1, N_PARAM)
        self.apertures = np.asarray(apertures, np.float64)
        self.before = [int(x) for x in before]
        self.after = [int(x) for x in after]
        self.media = list(media)          # dispersion curves (BaseDispersionCurve-like)
        self.detector_z = float(detector_z)
        if not 1 <= len(self.z) <= _abi.LENS_MAX_SURFACES:
            raise ValueError("1..%d surfaces" % _abi.LENS_MAX_SURFACES)
        self.zernike = list(zernike) if zernike is not None else [None] * len(self.z)
        if len(self.zernike) != len(self.z):
            raise ValueError("one zernike entry per surface")
        # per surface: unit radius, ANSI indices, (J, 28) expansion matrix
        self._zj, self._zT, self.zernike_radius = [], [], np.zeros(len(self.z))
        for s, zd in enumerate(self.zernike):
            if zd is None:
                self._zj.append([])
                self._zT.append(np.zeros((0, _abi.LENS_ZTERMS)))
                continue
            ru, terms = zd
            if not ru > 0:
                raise ValueError("Zernike unit radius must be positive")
            self.zernike_radius[s] = float(ru)
            self._zj.append([int(j) for j, _ in terms])
            self._zT.append(np.stack([zernike_monomials(j) for j, _ in terms]) if terms
                            else np.zeros((0, _abi.LENS_ZTERMS)))

    @property
    def zcoefs(self):
        """All Zernike coefficients, surface by surface, in term order."""
        out = [c for zd in self.zernike if zd is not None for _, c in zd[1]]
        return np.asarray(out, np.float64)

    @property
    def has_zernike(self):
        return any(len(j) for j in self._zj)

    def zernike_table(self, zcoefs=None):
        """(n_surfaces, 28) monomial coefficients of every surface's
        distortion for the flat coefficient vector zcoefs."""
        zc = self.zcoefs if zcoefs is None else np.asarray(zcoefs, np.float64).reshape(-1)
        out = np.zeros((self.n_surfaces, _abi.LENS_ZTERMS))
        k = 0
        for s, T in enumerate(self._zT):
            J = T.shape[0]
            if J:
                out[s] = zc[k:k + J] @ T
            k += J
        return out

    def zernike_grad(self, gmono):
        """Gradient with respect to the flat coefficient vector from the
        gradient with respect to the monomial table ((n_surfaces, 28))."""
        out = []
        for s, T in enumerate(self._zT):
            if T.shape[0]:
                out.append(T @ gmono[s])
        return np.concatenate(out) if out else np.zeros(0)

    @property
    def n_surfaces(self):
        return len(self.z)

    @classmethod
    def from_prescription(cls, presc, media, detector_z, zernike=None):
        return cls([p["z"] for p in presc],
                   [[p["c"], p["k"]] + list(p["A"]) for p in presc],
                   [p["aperture"] for p in presc], [p["before"] for p in presc], [p["after"] for p in presc],
                   media, detector_z, zernike)

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
            L.zernike_radius[s] = self.zernike_radius[s]
        return L

    def face_lists(self, params=None, zcoefs=None):
        """The lens as drop-in faces: [lens FaceList, detector FaceList]."""
        from .ctracer import FaceList
        from .cfaces import AsphericFace, CircularFace, DistortionFace
        from .cshapes import CircleShape
        from .cmaterials import FullDielectricDispersiveMaterial, OpaqueMaterial
        from .cdistortions import ZernikeDistortion
        P = self.params if params is None else np.asarray(params, np.float64).reshape(-1, N_PARAM)
        zc = self.zcoefs if zcoefs is None else np.asarray(zcoefs, np.float64).reshape(-1)
        faces = []
        k0 = 0
        for s in range(self.n_surfaces):
            c, k, A4, A6, A8, A10 = (float(x) for x in P[s])
            # inverted normals: the material's outside is the -z side (before)
            mat = FullDielectricDispersiveMaterial(dispersion_inside=self.media[self.after[s]],
                                                   dispersion_outside=self.media[self.before[s]])
            face = AsphericFace(shape=CircleShape(radius=float(self.apertures[s])), material=mat,
                                conic_const=k, z_height=float(self.z[s]), curvature=-1.0 / c,
                                A4=A4, A6=A6, A8=A8, A10=A10, invert_normals=True, atol=1e-16)
            js = self._zj[s]
            if js:
                zd = ZernikeDistortion([(j, float(v)) for j, v in zip(js, zc[k0:k0 + len(js)])],
                                       unit_radius=float(self.zernike_radius[s]))
                k0 += len(js)
                # the distorted normal is built with +z orientation (cfaces.pyx:2418-2431); the
                # base face's inverted normals point -z, so the orientation is flipped back
                face = DistortionFace(base_face=face, distortion=zd, material=mat, accuracy=1e-13, invert_normal=1)
            faces.append(face)
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


def _ztable(lens, zcoefs):
    """Device monomial table of the lens's distortions, or None."""
    import torch
    if not lens.has_zernike:
        return None
    return torch.from_numpy(lens.zernike_table(zcoefs).reshape(-1)).to("cuda")


def _eval(lens, d, params, zcoefs=None, objective=_abi.LENS_SPOT_RMS):
    """Moments (device, LENS_MOMENTS doubles) of the objective."""
    import torch
    lib = load()
    L = lens.struct(params)
    s = d.dev.struct()
    zt = _ztable(lens, zcoefs)
    m = torch.zeros(_abi.LENS_MOMENTS, dtype=torch.float64, device="cuda")
    check(lib.rpx_lens_eval(ctypes.byref(L), ctypes.c_void_p(zt.data_ptr() if zt is not None else 0), ctypes.byref(s),
                            d.rays.n_rays, ctypes.c_void_p(d.table.data_ptr()), d.n_wl, int(objective),
                            ctypes.c_void_p(m.data_ptr()), ctypes.c_void_p(d.ws.data_ptr()), d.ws.numel(),
                            stream_handle()))
    m._rpx_ztable = zt   # keeps the table alive while the stream uses it
    return m


def _value(m, objective):
    """Objective value (0-d tensor) from its moments."""
    import torch
    if objective == _abi.LENS_SPOT_RMS:
        return _sigma(m)
    return torch.sqrt(torch.clamp(m[1] / m[0], min=0.0))


def _lens_grad(lens, d, P, zcoefs, m, objective):
    """(d objective / d params (n_surfaces, 6), d objective / d zcoefs) as
    device tensors."""
    import torch
    lib = load()
    L = lens.struct(P)
    s = d.dev.struct()
    zt = _ztable(lens, zcoefs)
    S = lens.n_surfaces
    g = torch.empty(S * (N_PARAM + _abi.LENS_ZTERMS), dtype=torch.float64, device="cuda")
    check(lib.rpx_lens_grad(ctypes.byref(L), ctypes.c_void_p(zt.data_ptr() if zt is not None else 0), ctypes.byref(s),
                            d.rays.n_rays, ctypes.c_void_p(d.table.data_ptr()), d.n_wl, int(objective),
                            ctypes.c_void_p(m.data_ptr()), ctypes.c_void_p(g.data_ptr()),
                            ctypes.c_void_p(d.ws.data_ptr()), d.ws.numel(), stream_handle()))
    gp = g[:S * N_PARAM].view(S, N_PARAM)
    gz = g[S * N_PARAM:].view(S, _abi.LENS_ZTERMS).cpu().numpy()
    return gp, torch.from_numpy(lens.zernike_grad(gz))


def objective_and_grad(lens, rays, objective="spot", params=None, zcoefs=None):
    """Objective value and its gradients: (value, d/d params as an
    (n_surfaces, 6) array, d/d zcoefs as a flat array in lens.zcoefs
    order)."""
    obj = OBJECTIVES[objective]
    P = lens.params if params is None else np.asarray(params, np.float64).reshape(-1, N_PARAM)
    zc = lens.zcoefs if zcoefs is None else np.asarray(zcoefs, np.float64).reshape(-1)
    d = _Device(lens, rays)
    m = _eval(lens, d, P, zc, obj)
    gp, gz = _lens_grad(lens, d, P, zc, m, obj)
    return float(_value(m, obj)), gp.cpu().numpy(), gz.numpy()


def objective_moments(lens, rays, objective="spot", params=None, zcoefs=None):
    """The objective's moments (see include/rpx.h, rpx_lens_eval)."""
    obj = OBJECTIVES[objective]
    d = _Device(lens, rays)
    return _eval(lens, d, lens.params if params is None else params, zcoefs, obj).cpu().numpy()


def _moments(lens, d, params):
    """Spot moments through rpx_lens_spot, which ignores distortions."""
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
    return _eval(lens, d, lens.params if params is None else params)[:4].cpu().numpy()


def spot_rms_and_grad(lens, rays, params=None):
    """Spot RMS (mm) and d(spot RMS)/d(params) as a (n_surfaces, 6) array
    (a lens with distortions keeps its Zernike coefficients fixed)."""
    if lens.has_zernike:
        v, gp, _ = objective_and_grad(lens, rays, "spot", params)
        return v, gp
    P = lens.params if params is None else np.asarray(params, np.float64).reshape(-1, N_PARAM)
    d = _Device(lens, rays)
    m = _moments(lens, d, P)
    g = _grad(lens, d, P, m)
    return float(_sigma(m)), g.cpu().numpy().reshape(-1, N_PARAM)


def _grad(lens, d, P, m):
    """d(spot RMS)/d(params) (device, flat) through rpx_lens_spot_grad,
    which ignores distortions."""
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

    class LensObjective(torch.autograd.Function):
        """Spot RMS or wavefront RMS (0-d tensor) of `lens` for `rays` as a
        function of the (n_surfaces, 6) parameter tensor and the flat
        Zernike coefficient tensor (lens.zcoefs order)."""

        @staticmethod
        def forward(ctx, params, zcoefs, lens, rays, objective):
            require_gpu()
            obj = OBJECTIVES[objective]
            P = params.detach().to("cpu", torch.float64).numpy()
            Z = zcoefs.detach().to("cpu", torch.float64).numpy()
            d = _Device(lens, rays)
            m = _eval(lens, d, P, Z, obj)
            ctx.lens, ctx.d, ctx.P, ctx.Z, ctx.m, ctx.obj = lens, d, P, Z, m, obj
            ctx.pdev, ctx.pdtype, ctx.zdev, ctx.zdtype = params.device, params.dtype, zcoefs.device, zcoefs.dtype
            return _value(m, obj).to(params.device, params.dtype)

        @staticmethod
        def backward(ctx, gout):
            gp, gz = _lens_grad(ctx.lens, ctx.d, ctx.P, ctx.Z, ctx.m, ctx.obj)
            gp = gp.to(ctx.pdev, ctx.pdtype) * gout.to(ctx.pdev)
            gz = gz.to(ctx.zdev, ctx.zdtype) * gout.to(ctx.zdev)
            return gp, gz, None, None, None

    return SpotRMS, LensObjective


_FUNCS = None


def spot_rms(params, lens, rays):
    """Differentiable spot RMS: params is a (n_surfaces, 6) float64 tensor
    (a lens with distortions keeps its Zernike coefficients fixed)."""
    global _FUNCS
    if _FUNCS is None:
        _FUNCS = _function()
    if lens.has_zernike:
        import torch
        return _FUNCS[1].apply(params, torch.from_numpy(lens.zcoefs), lens, rays, "spot")
    return _FUNCS[0].apply(params, lens, rays)


def lens_objective(params, zcoefs, lens, rays, objective="spot"):
    """Differentiable spot RMS ("spot") or wavefront RMS ("wavefront"):
    params is an (n_surfaces, 6) float64 tensor, zcoefs the flat float64
    tensor of the lens's Zernike coefficients (lens.zcoefs order; empty
    without distortions)."""
    global _FUNCS
    if _FUNCS is None:
        _FUNCS = _function()
    return _FUNCS[1].apply(params, zcoefs, lens, rays, objective)
