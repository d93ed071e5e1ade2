"""Gausslet mode fit and E-field summation (drop-in for raypier.core.cfields).

Both run on the GPU (rpx_fields.hip); inputs and outputs are numpy arrays as
in the reference, and device tensors are accepted too.
"""
import ctypes

import numpy as np

from . import _abi
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


# ------------------------------------------------------ beam decomposition
def calc_mode_curvature(rx, ry, dx, dy, dx2, dy2, dxdy):
    """(A, B, C, x', y', z'): Gaussian mode curvature coefficients and local
    basis at each point from the wavefront's derivatives
    (cfields.pyx:231-293)."""
    torch = require_gpu()
    n = len(rx)
    ins = [_dev(np.ascontiguousarray(v, np.float64).reshape(-1)) for v in (dx, dy, dx2, dy2, dxdy)]
    A, B, C = (torch.empty(max(n, 1), dtype=torch.float64, device="cuda") for _ in range(3))
    X, Y, Z = (torch.empty((max(n, 1), 3), dtype=torch.float64, device="cuda") for _ in range(3))
    check(load().rpx_calc_mode_curvature(*[_ptr(t) for t in ins], n, _ptr(A), _ptr(B), _ptr(C), _ptr(X), _ptr(Y),
                                         _ptr(Z), stream_handle()))
    return tuple(t[:n].cpu().numpy() for t in (A, B, C, X, Y, Z))


def apply_mode_curvature(gc, A, B, C):
    """Add the curvature coefficients to the parabasal directions of a
    GaussletCollection (cfields.pyx:397-424)."""
    n = gc.n_rays
    if n == 0:
        return
    dA, dB, dC = (_dev(np.ascontiguousarray(v, np.float64).reshape(-1)[:n]) for v in (A, B, C))
    s = gc._ensure_device().struct()
    check(load().rpx_apply_mode_curvature(ctypes.byref(s), n, _ptr(dA), _ptr(dB), _ptr(dC), stream_handle()))
    gc._device_changed()


def build_interaction_matrix(rx, ry, tx, ty, A, B, C, x, y, z, wavelength, spacing, max_spacing, blending,
                             test_spacing):
    """(data, (i, j)) COO entries of the field of mode i at test point j
    for every pair closer than max_spacing (cfields.pyx:295-395)."""
    torch = require_gpu()
    lib = load()
    n, l = len(rx), len(tx)
    f = lambda v: _dev(np.ascontiguousarray(v, np.float64).reshape(-1))
    d_rx, d_ry, d_tx, d_ty, d_A, d_B, d_C = (f(v) for v in (rx, ry, tx, ty, A, B, C))
    d_x, d_y, d_z = (f(np.asarray(v).reshape(-1, 3)) for v in (x, y, z))
    # the reference's own output estimate, grown on demand
    cap = max(n * (1 + int((7 * max_spacing * max_spacing) / (test_spacing * test_spacing))), 1)
    ws = torch.empty(max(int(lib.rpx_trace_workspace_bytes(n * l)), 8), dtype=torch.uint8, device="cuda")
    while True:
        data = torch.empty(2 * cap, dtype=torch.float64, device="cuda")
        ij = torch.empty(2 * cap, dtype=torch.int64, device="cuda")
        k = ctypes.c_uint64(0)
        code = lib.rpx_build_interaction_matrix(_ptr(d_rx), _ptr(d_ry), n, _ptr(d_tx), _ptr(d_ty), l, _ptr(d_A),
                                                _ptr(d_B), _ptr(d_C), _ptr(d_x), _ptr(d_y), _ptr(d_z),
                                                float(wavelength), float(spacing), float(max_spacing),
                                                float(blending), _ptr(data), _ptr(ij), cap, ctypes.byref(k),
                                                _ptr(ws), ws.numel(), stream_handle())
        if code == _abi.RPX_E_CAPACITY:
            cap = int(k.value)
            continue
        check(code)
        break
    m = int(k.value)
    vals = data[: 2 * m].cpu().numpy().view(np.complex128)
    idx = ij.cpu().numpy().view(np.uint64)
    return vals.copy(), (idx[:m].copy(), idx[cap: cap + m].copy())
