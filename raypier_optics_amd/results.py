"""Results computed over traced generations on the GPU.

`optical_path_lengths` / `mean_optical_path_length` are the parent-chain walk
of MeanOpticalPathLength._calc_result (results.py:112-122): for every ray of
the last generation that ended on the target face, the sum of length * Re(n)
over the ray and all its ancestors.  The walk runs on the device over the
generations where the trace left them (rpx_path_lengths); only the per-ray
totals come back, and their mean is taken with numpy as the reference does.
"""
import ctypes

import numpy as np

from . import _abi
from ._lib import load, check, require_gpu, stream_handle


def generation_table(traced_rays):
    """ctypes array of rpx_generation_t for a list of traced collections
    (kept alive by the returned device buffers)."""
    gens = (_abi.Generation * len(traced_rays))()
    devs = []
    for k, rc in enumerate(traced_rays):
        dev = rc._ensure_device()
        devs.append(dev)
        gens[k].rays = dev.struct()
        gens[k].n_rays = rc.n_rays
    return gens, devs


def optical_path_lengths(traced_rays, target_face):
    """float64 array: the optical path of each ray of traced_rays[-1] whose
    end_face_idx is target_face.idx (or the index itself), in ray order."""
    torch = require_gpu()
    lib = load()
    if len(traced_rays) == 0:
        raise IndexError("no traced generations")
    idx = int(getattr(target_face, "idx", target_face))
    n_last = traced_rays[-1].n_rays
    if n_last == 0:
        return np.zeros(0, np.float64)
    if not 0 <= idx <= 0xFFFFFFFF:
        return np.zeros(0, np.float64)   # an end_face_idx is never negative
    gens, devs = generation_table(traced_rays)
    ws = torch.empty(int(lib.rpx_lineage_workspace_bytes(len(traced_rays), n_last)), dtype=torch.uint8,
                     device="cuda")
    totals = torch.empty(n_last, dtype=torch.float64, device="cuda")
    n_sel = ctypes.c_uint64(0)
    check(lib.rpx_path_lengths(gens, len(traced_rays), idx, ctypes.c_void_p(totals.data_ptr()), n_last,
                               ctypes.byref(n_sel), ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream_handle()))
    return totals[: n_sel.value].cpu().numpy()


def mean_optical_path_length(traced_rays, target_face):
    """MeanOpticalPathLength.result (results.py:112-122)."""
    return optical_path_lengths(traced_rays, target_face).mean()
