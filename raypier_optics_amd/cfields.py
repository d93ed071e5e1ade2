"""Gausslet mode fit and E-field summation (drop-in for raypier.core.cfields).

Both run on the GPU (rpx_fields.hip); inputs and outputs are numpy arrays as
in the reference, and device tensors are accepted too.
"""
import ctypes

import numpy as np

from ._lib import load, check, require_gpu, stream_handle


def _dev(a, dtype=np.float64):
    torch = require_gpu()
    if hasattr(a, "data_ptr"):
        return a.to(device="cuda", dtype=torch.float64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to("cuda")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def evaluate_modes(x, y, dx, dy, blending=1.0):
    """Per-row closed-form least-squares fit of the astigmatic Gaussian mode
    coefficients (cfields.pyx:156-228); returns complex (N, 3)."""
    torch = require_gpu()
    X, Y, DX, DY = (_dev(np.asarray(v, np.float64)) for v in (x, y, dx, dy))
    n, row = X.shape
    out = torch.empty((n, 3, 2), dtype=torch.float64, device="cuda")
    check(load().rpx_evaluate_modes(_ptr(X), _ptr(Y), _ptr(DX), _ptr(DY), n, row, float(blending), _ptr(out),
                                    stream_handle()))
    o = out.cpu().numpy()
    return o[..., 0] + 1j * o[..., 1]


def _modes_device(modes):
    torch = require_gpu()
    m = np.ascontiguousarray(modes, dtype=np.complex128)
    return torch.from_numpy(m.view(np.float64).copy()).to("cuda")


def sum_gaussian_modes(rays, modes, wavelengths, points, time_ps=0.0, precision=0, as_tensor=False):
    """E-field at `points` from base rays and their mode coefficients
    (cfields.pyx:52-115); returns complex (Npt, 3)."""
    torch = require_gpu()
    dev = rays._ensure_device()
    n = rays.n_rays
    md = modes if hasattr(modes, "data_ptr") else _modes_device(modes)
    wl = _dev(np.asarray(wavelengths, np.float64)) if not hasattr(wavelengths, "data_ptr") else wavelengths
    pts = _dev(points).reshape(-1, 3).contiguous()
    npt = pts.shape[0]
    out = torch.empty((npt, 3, 2), dtype=torch.float64, device="cuda")
    s = dev.struct()
    check(load().rpx_sum_gaussian_modes(ctypes.byref(s), n, _ptr(md), _ptr(wl), wl.numel(), _ptr(pts), npt,
                                        float(time_ps), _ptr(out), int(precision), stream_handle()))
    if as_tensor:
        return torch.view_as_complex(out)
    o = out.cpu().numpy()
    return o[..., 0] + 1j * o[..., 1]


def gausslet_modes(gausslets, blending=1.0):
    """evaluate_neighbours_gc + evaluate_modes fused on the GPU; returns the
    (N, 3) complex modes as a device float64 tensor of shape (N, 6)."""
    torch = require_gpu()
    dev = gausslets._ensure_device()
    n = gausslets.n_rays
    out = torch.empty((max(n, 1), 6), dtype=torch.float64, device="cuda")
    s = dev.struct()
    check(load().rpx_gausslet_modes(ctypes.byref(s), n, float(blending), _ptr(out), stream_handle()))
    return out[:n]
