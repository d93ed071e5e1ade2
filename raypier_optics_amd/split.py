"""FP32-split trace form.

Ray generations are stored in HBM in float32 (vectors, refractive index,
amplitudes, length) with float64 phase and accumulated optical path: 112
bytes per ray against 188.  The FP32-split kernels compute root solves,
normals and child directions in float64 from the float32 state and the
polarisation amplitudes in float32 (include/rpx.h, rpx_trace_segment_split).
The reference traces in float64 only; this form has no counterpart there.

`trace_rays_split` takes the same arguments as `tracer.trace_rays`
(core/tracer.py:9-47) and returns (generations, all_faces), where each
generation is a `SplitRayCollection`; `to_collection()` converts one back to
a float64 RayCollection.
"""
import ctypes

import numpy as np

from . import _abi
from ._lib import load, check, require_gpu, stream_handle
from .ctracer import RayCollection, _TraceContext, _wavelength_table
from .tracer import _prepare


class SplitRays(object):
    """HBM-resident FP32-split SoA buffer for up to `capacity` rays: one
    allocation with the 2 float64 rows, then 19 float32 rows, then 5 uint32
    rows of `capacity` elements each."""

    def __init__(self, capacity):
        torch = require_gpu()
        self.capacity = cap = int(max(capacity, 1))
        nbytes = cap * (_abi.SPLIT_F64 * 8 + _abi.SPLIT_F32 * 4 + _abi.N_U32 * 4)
        self.buf = torch.empty(nbytes + 256, dtype=torch.uint8, device="cuda")

    def _offsets(self):
        cap = self.capacity
        o64 = 0
        o32 = o64 + _abi.SPLIT_F64 * 8 * cap
        ou = o32 + _abi.SPLIT_F32 * 4 * cap
        return o64, o32, ou

    def f64(self):
        import torch
        o64, o32, _ = self._offsets()
        return self.buf[o64:o32].view(torch.float64).view(_abi.SPLIT_F64, self.capacity)

    def f32(self):
        import torch
        _, o32, ou = self._offsets()
        return self.buf[o32:ou].view(torch.float32).view(_abi.SPLIT_F32, self.capacity)

    def u32(self):
        import torch
        _, _, ou = self._offsets()
        return self.buf[ou: ou + _abi.N_U32 * 4 * self.capacity].view(torch.int32).view(_abi.N_U32, self.capacity)

    def struct(self):
        base = self.buf.data_ptr()
        o64, o32, ou = self._offsets()
        return _abi.RaysSplit(base + o32, base + o64, base + ou, self.capacity)


class SplitRaysView(SplitRays):
    """Rays [offset, offset + n) of a pool SplitRays (same buffer and field
    stride): a generation written into a pool, without a copy.  Holds a
    reference to the pool."""

    def __init__(self, pool, offset, n):
        self.pool = pool
        self.offset = int(offset)
        self.capacity = int(n)
        self.buf = pool.buf

    def f64(self):
        return self.pool.f64()[:, self.offset: self.offset + self.capacity]

    def f32(self):
        return self.pool.f32()[:, self.offset: self.offset + self.capacity]

    def u32(self):
        return self.pool.u32()[:, self.offset: self.offset + self.capacity]

    def struct(self):
        s = self.pool.struct()
        o = self.offset
        return _abi.RaysSplit(s.f32 + 4 * o, s.f64 + 8 * o, s.u32 + 4 * o, self.pool.capacity)


_F = {name: k for k, name in enumerate(_abi.F64_FIELDS[: _abi.SPLIT_F32])}
_U = {name: k for k, name in enumerate(_abi.U32_FIELDS)}
_VEC = {"origin": "ox", "direction": "dx", "normal": "nx", "E_vector": "ex"}


class SplitRayCollection(object):
    """One FP32-split generation of rays (device-resident)."""

    def __init__(self, dev, n, wavelengths, wl_dev):
        self.dev = dev
        self._n = int(n)
        self._wavelengths = wavelengths
        self._wl_dev = wl_dev
        self.parent = None

    @property
    def n_rays(self):
        return self._n

    def __len__(self):
        return self._n

    @property
    def wavelengths(self):
        return np.asarray(self._wavelengths)

    def _device_wavelengths(self):
        return self._wl_dev

    def reset_length(self, max_length=float("inf")):
        """Set the length of every ray to max_length (rounded to float32)."""
        self.dev.f32()[_F["length"], : self._n].fill_(float(max_length))

    # float32 / uint32 field views (torch, on the device) ----------------
    def field(self, name):
        """Device view of one field: a float32 vector field as (n, 3), a
        float32 scalar as (n,), a uint32 field as int32 (n,), phase and
        accumulated_path as float64."""
        n = self._n
        if name in _VEC:
            k = _F[_VEC[name]]
            return self.dev.f32()[k: k + 3, :n].t()
        if name in ("phase", "accumulated_path"):
            return self.dev.f64()[0 if name == "phase" else 1, :n]
        if name in _U:
            return self.dev.u32()[_U[name], :n]
        return self.dev.f32()[_F[name], :n]

    def to_collection(self):
        """Float64 RayCollection with the same rays (fields widened)."""
        from ._device import DeviceRays
        out = DeviceRays(max(self._n, 1))
        s_in, s_out = self.dev.struct(), out.struct()
        check(load().rpx_rays_from_split(ctypes.byref(s_in), self._n, ctypes.byref(s_out), stream_handle()))
        rc = RayCollection.from_device(out, self._n)
        rc.wavelengths = self._wavelengths
        return rc

    @classmethod
    def from_collection(cls, rays):
        """FP32-split copy of a float64 RayCollection (fields rounded)."""
        n = rays.n_rays
        src = rays._ensure_device()
        dev = SplitRays(n)
        s_in, s_out = src.struct(), dev.struct()
        check(load().rpx_rays_to_split(ctypes.byref(s_in), n, ctypes.byref(s_out), stream_handle()))
        return cls(dev, n, rays.wavelengths, rays._device_wavelengths())


def _split_generation(ctx, rays, max_length, one_list=-1, one_face=-1):
    lib = load()
    n = rays.n_rays
    cap = max(ctx.max_children * n, 1)
    out = SplitRays(cap)
    ws = ctx.workspace(n)
    s_in, s_out = rays.dev.struct(), out.struct()
    n_out = ctypes.c_uint64(0)
    check(lib.rpx_trace_segment_split(ctx.handle, one_list, one_face, ctypes.byref(s_in), n, float(max_length),
                                      ctypes.c_void_p(rays._wl_dev.data_ptr()), len(rays._wavelengths),
                                      ctypes.byref(s_out), ctypes.byref(n_out), ctypes.c_void_p(ctx.hits.data_ptr()),
                                      ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream_handle()))
    child = SplitRayCollection(out, n_out.value, rays._wavelengths, rays._wl_dev)
    child.parent = rays
    return child


def _split_chain(ctx, rays, max_length, limit):
    """The FP32-split generations on the device (rpx_trace_rays_split), with
    the pool and resume bookkeeping of tracer._trace_chain."""
    from .tracer import _pool_rays, _alloc_pool
    lib = load()
    traced = [rays]
    current = rays
    remaining = limit
    hits = (ctypes.c_uint32 * max(ctx.tables.n_global_faces, 1))()
    while remaining > 0 and current.n_rays > 0:
        n = current.n_rays
        pool = _alloc_pool(SplitRays, _pool_rays(n, ctx.max_children, remaining),
                           max(ctx.max_children, 1) * n if remaining > 1 else 1)
        ws_bytes = int(lib.rpx_chain_workspace_bytes(max(n, pool.capacity), remaining))
        ws = ctx.chain_workspace(ws_bytes)
        counts = (ctypes.c_uint64 * (remaining + 1))()
        offsets = (ctypes.c_uint64 * (remaining + 1))()
        n_traced = ctypes.c_int32(0)
        s_in, s_pool = current.dev.struct(), pool.struct()
        code = lib.rpx_trace_rays_split(ctx.handle, ctypes.byref(s_in), n, float(max_length),
                                        ctypes.c_void_p(current._wl_dev.data_ptr()), len(current._wavelengths),
                                        remaining, ctypes.byref(s_pool), counts, offsets, ctypes.byref(n_traced),
                                        hits, ctypes.c_void_p(ws.data_ptr()), ws_bytes, stream_handle())
        if code != _abi.RPX_E_CAPACITY:
            check(code)
        k_traced = n_traced.value
        last = k_traced + 1 if code == _abi.RPX_E_CAPACITY else k_traced
        parent = current
        for k in range(1, last):
            gen = SplitRayCollection(SplitRaysView(pool, offsets[k], counts[k]), counts[k], current._wavelengths,
                                     current._wl_dev)
            gen.parent = parent
            traced.append(gen)
            parent = gen
        remaining -= k_traced
        if code != _abi.RPX_E_CAPACITY:
            break
        current = traced[-1]
        if k_traced == 0:
            raise RuntimeError("device-resident trace made no progress")
    for g, h in enumerate(hits):
        if h and g < len(ctx.all_faces):
            ctx.all_faces[g].count += int(h)
    return traced


def trace_rays_split(input_rays, face_lists, recursion_limit=100, max_length=100.0):
    """FP32-split non-sequential trace of a RayCollection; generation
    bookkeeping as tracer.trace_rays (core/tracer.py:9-47).  The generations
    run device-resident (rpx_trace_rays_split), which writes every input
    ray's length itself."""
    all_faces = _prepare(input_rays, face_lists, max_length, reset=False)
    face_sets = list(face_lists)
    for fs in face_sets:
        fs.sync_transforms()
    ctx = _TraceContext(face_sets, all_faces)
    _wavelength_table(input_rays, ctx)
    rays = input_rays if isinstance(input_rays, SplitRayCollection) else SplitRayCollection.from_collection(input_rays)
    if rays.n_rays == 0 or recursion_limit <= 0:
        return [], all_faces
    return _split_chain(ctx, rays, max_length, recursion_limit), all_faces


def trace_ray_sequence_split(input_rays, face_sequence, recursion_limit=100, max_length=100.0):
    """FP32-split sequential trace through (FaceList, face_idx) pairs, with
    the bookkeeping of tracer.trace_ray_sequence (core/tracer.py:50-99)."""
    input_rays.reset_length(max_length)
    face_lists = [fl for fl, fidx in face_sequence]
    all_faces = _prepare(input_rays, face_lists, max_length)
    rays = SplitRayCollection.from_collection(input_rays)
    traced = [rays]
    contexts = {}
    count = 0
    for fl, fidx in face_sequence:
        fl.sync_transforms()
        key = id(fl)
        if key not in contexts:
            contexts[key] = _TraceContext([fl], all_faces)
        ctx = contexts[key]
        rays = _split_generation(ctx, rays, max_length, 0, fidx)
        ctx.flush_counts()
        if (count > recursion_limit) or (rays.n_rays == 0):
            break
        traced.append(rays)
        count += 1
    return traced, all_faces
