"""Loader and ctypes bindings for librpx.so (the gfx950 HIP library).

The library is loaded from the package directory.  There is no fallback: if
it is missing or no GPU is usable, tracing raises.
"""
import ctypes
import os

from . import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librpx.so")

_lib = None

c_void_p = ctypes.c_void_p
c_u64 = ctypes.c_uint64
c_i32 = ctypes.c_int32
c_u32 = ctypes.c_uint32
c_f64 = ctypes.c_double
c_f32 = ctypes.c_float
P = ctypes.POINTER

_SIGNATURES = {
    "rpx_abi_version": (c_i32, []),
    "rpx_last_error": (ctypes.c_char_p, []),
    "rpx_struct_size": (c_i32, [ctypes.c_char_p]),
    "rpx_scene_create": (c_i32, [P(_abi.SceneDesc), P(c_void_p)]),
    "rpx_scene_destroy": (c_i32, [c_void_p]),
    "rpx_scene_max_children": (c_i32, [c_void_p]),
    "rpx_trace_workspace_bytes": (c_u64, [c_u64]),
    "rpx_trace_segment": (c_i32, [c_void_p, P(_abi.Rays), c_u64, c_f32, c_void_p, c_u64, P(_abi.Rays),
                                  P(c_u64), c_void_p, c_void_p, c_u64, c_void_p]),
    "rpx_chain_workspace_bytes": (c_u64, [c_u64, c_i32]),
    "rpx_trace_rays": (c_i32, [c_void_p, P(_abi.Rays), c_u64, c_f32, c_void_p, c_u64, c_i32, P(_abi.Rays),
                               c_void_p, c_void_p, P(c_i32), c_void_p, c_void_p, c_u64, c_void_p]),
    "rpx_trace_gausslets": (c_i32, [c_void_p, P(_abi.Rays), c_u64, c_f64, c_void_p, c_u64, c_i32, P(_abi.Rays),
                               c_void_p, c_void_p, P(c_i32), c_void_p, c_void_p, c_u64, c_void_p]),
    "rpx_trace_one_face_segment": (c_i32, [c_void_p, c_i32, c_i32, P(_abi.Rays), c_u64, c_f32, c_void_p, c_u64,
                                           P(_abi.Rays), P(c_u64), c_void_p, c_void_p, c_u64, c_void_p]),
    "rpx_trace_gausslet": (c_i32, [c_void_p, c_i32, c_i32, P(_abi.Rays), c_u64, c_f64, c_void_p, c_u64,
                                   P(_abi.Rays), P(c_u64), c_void_p, c_void_p, c_u64, c_void_p]),
    "rpx_select_intersections": (c_i32, [c_void_p, c_i32, P(_abi.Rays), c_u64, c_i32, c_void_p, c_u64,
                                         P(_abi.Rays), c_u64, P(c_u64), c_void_p, c_u64, c_void_p]),
    "rpx_rays_to_split": (c_i32, [P(_abi.Rays), c_u64, P(_abi.RaysSplit), c_void_p]),
    "rpx_rays_from_split": (c_i32, [P(_abi.RaysSplit), c_u64, P(_abi.Rays), c_void_p]),
    "rpx_trace_segment_split": (c_i32, [c_void_p, c_i32, c_i32, P(_abi.RaysSplit), c_u64, c_f32, c_void_p, c_u64,
                                        P(_abi.RaysSplit), P(c_u64), c_void_p, c_void_p, c_u64, c_void_p]),
    "rpx_trace_rays_split": (c_i32, [c_void_p, P(_abi.RaysSplit), c_u64, c_f32, c_void_p, c_u64, c_i32,
                                     P(_abi.RaysSplit), c_void_p, c_void_p, P(c_i32), c_void_p, c_void_p, c_u64,
                                     c_void_p]),
    "rpx_tile_order_fallbacks": (c_u64, []),
    "rpx_debug_lookback_ticks": (c_i32, [c_u64]),
    "rpx_server_create": (c_i32, [c_u64, c_u32, c_u64, c_void_p, P(c_void_p)]),
    "rpx_server_destroy": (c_i32, [c_void_p]),
    "rpx_server_submit": (c_i32, [c_void_p, P(_abi.JobDesc), P(c_u64)]),
    "rpx_server_step": (c_i32, [c_void_p, c_void_p, c_u32, P(c_u32)]),
    "rpx_server_counts": (c_i32, [c_void_p, P(c_u32), P(c_u32)]),
    "rpx_server_stats": (c_i32, [c_void_p, c_void_p, c_void_p]),
    "rpx_server_clock": (c_f64, []),
    "rpx_server_set_generations_per_step": (c_i32, [c_void_p, c_u32]),
    "rpx_server_set_graph_launch": (c_i32, [c_void_p, c_i32]),
    "rpx_server_fail_allocations_after": (c_i32, [c_void_p, ctypes.c_int64]),
    "rpx_lens_workspace_bytes": (c_u64, [c_u64, c_i32]),
    "rpx_lens_eval": (c_i32, [P(_abi.Lens), c_void_p, P(_abi.Rays), c_u64, c_void_p, c_u64, c_i32, c_void_p,
                              c_void_p, c_u64, c_void_p]),
    "rpx_lens_grad": (c_i32, [P(_abi.Lens), c_void_p, P(_abi.Rays), c_u64, c_void_p, c_u64, c_i32, c_void_p,
                              c_void_p, c_void_p, c_u64, c_void_p]),
    "rpx_lens_spot": (c_i32, [P(_abi.Lens), P(_abi.Rays), c_u64, c_void_p, c_u64, c_void_p, c_void_p, c_u64,
                              c_void_p]),
    "rpx_lens_spot_grad": (c_i32, [P(_abi.Lens), P(_abi.Rays), c_u64, c_void_p, c_u64, c_void_p, c_void_p,
                                   c_void_p, c_u64, c_void_p]),
    "rpx_lineage_workspace_bytes": (c_u64, [c_i32, c_u64]),
    "rpx_path_lengths": (c_i32, [P(_abi.Generation), c_i32, c_u32, c_void_p, c_u64, P(c_u64), c_void_p, c_u64,
                                 c_void_p]),
    "rpx_face_histogram": (c_i32, [P(_abi.Generation), c_i32, c_u32, c_void_p, c_void_p, c_u64, c_void_p]),
    "rpx_neighbours_workspace_bytes": (c_u64, [c_u64]),
    "rpx_propagate_neighbours": (c_i32, [P(_abi.Rays), c_u64, c_void_p, c_u64, c_i32, c_void_p, c_void_p, c_u64,
                                         c_void_p]),
    "rpx_lattice_workspace_bytes": (c_u64, [c_i32]),
    "rpx_lattice_select": (c_i32, [P(_abi.Lattice), P(c_u64), P(c_f64), c_void_p, c_u64, c_void_p]),
    "rpx_lattice_emit": (c_i32, [P(_abi.Lattice), P(_abi.Rays), c_void_p, c_void_p, c_u64, c_void_p]),
    "rpx_config_parabasal_rays": (c_i32, [P(_abi.Rays), c_u64, c_void_p, c_u64, c_f64, c_f64, P(c_f64), c_void_p,
                                          c_u64, c_void_p]),
    "rpx_project_to_plane": (c_i32, [P(_abi.Rays), c_u64, P(c_f64), P(c_f64), c_void_p]),
    "rpx_lagrange_invariant": (c_i32, [P(_abi.Rays), c_u64, c_void_p, c_u64, c_void_p, c_void_p, c_u64, c_void_p]),
    "rpx_project_to_sphere": (c_i32, [P(_abi.Rays), c_u64, P(c_f64), c_f64, P(_abi.Rays), P(c_u64), c_void_p, c_u64,
                                      c_void_p]),
    "rpx_ray_field_modes": (c_i32, [P(_abi.Rays), c_u64, c_void_p, c_f64, P(_abi.Rays), c_void_p, c_void_p,
                                    P(c_u64), c_void_p, c_u64, c_void_p]),
    "rpx_capture_gausslets": (c_i32, [P(_abi.Rays), c_u64, P(c_i32), c_i32, c_void_p, c_u64, P(_abi.Rays), c_u64,
                                      P(c_u64), c_void_p, c_u64, c_void_p]),
    "rpx_calc_mode_curvature": (c_i32, [c_void_p] * 5 + [c_u64] + [c_void_p] * 7),
    "rpx_apply_mode_curvature": (c_i32, [P(_abi.Rays), c_u64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "rpx_build_interaction_matrix": (c_i32, [c_void_p, c_void_p, c_u64, c_void_p, c_void_p, c_u64] + [c_void_p] * 6 +
                                     [c_f64] * 4 + [c_void_p, c_void_p, c_u64, P(c_u64), c_void_p, c_u64, c_void_p]),
    "rpx_timing_enable": (c_i32, [c_i32]),
    "rpx_timing_read": (c_i32, [c_void_p, c_void_p, c_void_p]),
    "rpx_query_face": (c_i32, [c_void_p, c_i32, c_i32, c_void_p, c_void_p, c_u64, c_i32, c_void_p, c_void_p]),
    "rpx_query_shape": (c_i32, [c_void_p, c_i32, c_void_p, c_u64, c_void_p, c_void_p]),
    "rpx_query_zernike": (c_i32, [c_void_p, c_i32, c_void_p, c_u64, c_void_p, c_void_p, c_void_p]),
    "rpx_eval_child_rays": (c_i32, [c_void_p, c_i32, P(_abi.Rays), c_u64, c_u32, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_u64, P(_abi.Rays), c_void_p, c_void_p]),
    "rpx_pack_rays": (c_i32, [P(_abi.Rays), c_u64, c_void_p, c_i32, c_void_p]),
    "rpx_unpack_rays": (c_i32, [c_void_p, c_u64, P(_abi.Rays), c_i32, c_void_p]),
    "rpx_gausslet_modes": (c_i32, [P(_abi.Rays), c_u64, c_f64, c_void_p, c_void_p]),
    "rpx_evaluate_modes": (c_i32, [c_void_p, c_void_p, c_void_p, c_void_p, c_u64, c_i32, c_f64, c_void_p,
                                   c_void_p]),
    "rpx_sum_gaussian_modes": (c_i32, [P(_abi.Rays), c_u64, c_void_p, c_void_p, c_u64, c_void_p, c_u64, c_f64,
                                       c_void_p, c_i32, c_void_p]),
    "rpx_field_noalloc_max_rays": (c_u64, []),
    "rpx_sum_gaussian_modes_noalloc": (c_i32, [P(_abi.Rays), c_u64, c_void_p, c_void_p, c_u64, c_void_p, c_u64,
                                               c_f64, c_void_p, c_i32, c_void_p]),
}

EXPORTED = sorted(_SIGNATURES)


class RpxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("librpx error %d: %s" % (code, msg))
        self.code = code


def load():
    """Load librpx.so (raises OSError with the path if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError("librpx.so is not built at %s; run raypier_optics_amd.build_lib.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.rpx_abi_version() != _abi.ABI_VERSION:
        raise OSError("librpx ABI version mismatch")
    for name, size in _abi.STRUCT_SIZES.items():
        got = lib.rpx_struct_size(name.encode())
        if got != size:
            raise OSError("struct %s: library has %d bytes, binding %d" % (name, got, size))
    _lib = lib
    return lib


def check(code, lib=None):
    if code != _abi.RPX_OK:
        lib = lib or load()
        raise RpxError(code, lib.rpx_last_error().decode(errors="replace"))


def require_gpu():
    """Return the torch module with a usable GPU, or raise."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("raypier_optics_amd traces on the GPU: no ROCm device is available")
    return torch


def stream_handle():
    torch = require_gpu()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
