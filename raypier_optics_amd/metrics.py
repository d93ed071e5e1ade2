"""Spot statistics on traced generations (float64 or FP32-split).

A detector hit is a ray of any generation whose end_face_idx is the
detector's face index; its point is origin + direction * length, evaluated
in float64.  `spot_rms_by_wavelength` bins the hits by wavelength and
returns, per bin, the RMS distance of the hit points from their centroid
(the spot RMS radius) -- the figure in which the FP32-split form's error
against float64 is stated.

`strehl_ratio` is the Strehl ratio at a reference point P of the rays that
reach the detector from one point source at one wavelength: the intensity at
P of the coherent sum of the rays, each with unit amplitude and the phase of
its optical path to P, over the same sum with every path equal,
|mean(exp(i phi))|^2 (uniform pupil).  A ray's path to P is its optical path
to its last origin Q (accumulated_path, in mm, plus its phase in rad) plus
n (P - Q).d, the path to the foot of the perpendicular from P: to first order
in the aberration, the wavefront error against the sphere centred on P.
"""
import numpy as np


def _fields(gen):
    """(origin, direction, length, end_face_idx, wavelength_idx) of one
    generation as float64 / int64 torch tensors on the device."""
    import torch
    if hasattr(gen, "field"):   # SplitRayCollection
        return (gen.field("origin").double(), gen.field("direction").double(), gen.field("length").double(),
                gen.field("end_face_idx").long() & 0xFFFFFFFF, gen.field("wavelength_idx").long() & 0xFFFFFFFF)
    from . import _abi
    dev = gen._ensure_device()
    n = gen.n_rays
    f = dev.f64()
    u = dev.u32()
    k = _abi.F64_FIELDS.index
    origin = f[k("ox"): k("ox") + 3, :n].t()
    direction = f[k("dx"): k("dx") + 3, :n].t()
    length = f[k("length"), :n]
    end = u[_abi.U32_FIELDS.index("end_face_idx"), :n].long() & 0xFFFFFFFF
    wl = u[_abi.U32_FIELDS.index("wavelength_idx"), :n].long() & 0xFFFFFFFF
    return origin, direction, length, end, wl


def detector_hits(traced, face_idx):
    """Hit points (n, 3) float64 and wavelength indices (n,) of every ray of
    `traced` (a list of generations) that ended on face `face_idx`."""
    import torch
    pts, wls = [], []
    for gen in traced:
        if gen.n_rays == 0:
            continue
        o, d, L, end, wl = _fields(gen)
        m = end == int(face_idx)
        pts.append(o[m] + d[m] * L[m, None])
        wls.append(wl[m])
    if not pts:
        return torch.zeros((0, 3), dtype=torch.float64), torch.zeros(0, dtype=torch.int64)
    return torch.cat(pts), torch.cat(wls)


def spot_rms_by_wavelength(traced, face_idx, wavelengths, edges):
    """Spot RMS radius and hit count per wavelength bin [edges[k], edges[k+1])."""
    import torch
    pts, wl_idx = detector_hits(traced, face_idx)
    wl = torch.as_tensor(np.asarray(wavelengths, np.float64), device=pts.device)[wl_idx]
    out_rms, out_n = [], []
    for lo, hi in zip(edges[:-1], edges[1:]):
        m = (wl >= lo) & (wl < hi)
        p = pts[m]
        n = int(p.shape[0])
        out_n.append(n)
        if n == 0:
            out_rms.append(float("nan"))
            continue
        c = p.mean(dim=0)
        out_rms.append(float(torch.sqrt(((p - c) ** 2).sum(dim=1).mean())))
    return np.array(out_rms), np.array(out_n)


def _path_fields(gen):
    """(origin, direction, accumulated_path, phase, n_re, end_face_idx) of
    one generation as float64 / int64 torch tensors on the device."""
    if hasattr(gen, "field"):   # SplitRayCollection
        return (gen.field("origin").double(), gen.field("direction").double(),
                gen.field("accumulated_path").double(), gen.field("phase").double(),
                gen.field("n_re").double(), gen.field("end_face_idx").long() & 0xFFFFFFFF)
    from . import _abi
    dev = gen._ensure_device()
    n = gen.n_rays
    f = dev.f64()
    k = _abi.F64_FIELDS.index
    return (f[k("ox"): k("ox") + 3, :n].t(), f[k("dx"): k("dx") + 3, :n].t(), f[k("accum"), :n],
            f[k("phase"), :n], f[k("n_re"), :n],
            dev.u32()[_abi.U32_FIELDS.index("end_face_idx"), :n].long() & 0xFFFFFFFF)


def detector_paths(traced, face_idx):
    """(Q, d, accumulated_path, phase, n_re) of the rays of `traced` that
    ended on face `face_idx`, concatenated over generations."""
    import torch
    parts = []
    for gen in traced:
        if gen.n_rays == 0:
            continue
        o, d, acc, ph, n, end = _path_fields(gen)
        m = end == int(face_idx)
        parts.append((o[m], d[m], acc[m], ph[m], n[m]))
    if not parts:
        z = torch.zeros(0, dtype=torch.float64)
        return torch.zeros((0, 3), dtype=torch.float64), torch.zeros((0, 3), dtype=torch.float64), z, z, z
    return tuple(torch.cat([p[i] for p in parts]) for i in range(5))


def strehl_ratio(paths, wavelength_um, point):
    """Strehl ratio at `point` (mm) of detector_paths() output for one
    wavelength (um)."""
    import torch
    Q, d, acc, ph, n = paths
    if Q.shape[0] == 0:
        return float("nan")
    P = torch.as_tensor(np.asarray(point, np.float64), device=Q.device)
    W = acc + n * ((P - Q) * d).sum(dim=1)
    phi = ph + (W - W[0]) * (2 * np.pi / (wavelength_um * 1e-3))
    c, s = torch.cos(phi).mean(), torch.sin(phi).mean()
    return float(c * c + s * s)


def best_focus_strehl(paths, wavelength_um, center, span_mm=0.2, steps=81):
    """(point, Strehl) maximising the Strehl ratio over points center + (0,
    0, dz), |dz| <= span_mm (center: e.g. the centroid of the detector
    hits)."""
    best = (None, -1.0)
    for dz in np.linspace(-span_mm, span_mm, steps):
        p = (float(center[0]), float(center[1]), float(center[2]) + dz)
        s = strehl_ratio(paths, wavelength_um, p)
        if s > best[1]:
            best = (p, s)
    return best
