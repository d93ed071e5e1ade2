"""Spot statistics on traced generations (float64 or FP32-split).

A detector hit is a ray of any generation whose end_face_idx is the
detector's face index; its point is origin + direction * length, evaluated
in float64.  `spot_rms_by_wavelength` bins the hits by wavelength and
returns, per bin, the RMS distance of the hit points from their centroid
(the spot RMS radius) -- the figure in which the FP32-split form's error
against float64 is stated.
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
