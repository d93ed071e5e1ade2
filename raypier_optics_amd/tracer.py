"""Generation drivers (drop-in for raypier.core.tracer)."""
from itertools import chain

import numpy

import ctypes

from . import _abi
from ._lib import load, check, require_gpu, stream_handle
from .ctracer import RayCollection, GaussletCollection, _TraceContext, _generation, _wavelength_table


def _prepare(input_rays, face_lists, max_length, reset=True):
    if reset:
        input_rays.reset_length(max_length)
    wavelengths = numpy.asarray(input_rays.wavelengths)
    all_faces = list(chain(*(fs.faces for fs in face_lists)))
    for i, f in enumerate(all_faces):
        f.idx = i
        f.count = 0
        f.update()
        f.material.wavelengths = wavelengths
        f.max_length = max_length
    return all_faces


# the device-resident chain is sized for at most this many generations ahead
_CHAIN_GENERATIONS = 16


_DEVICE_BYTES = {}


def _device_bytes(torch):
    """Total memory of the current device (looked up once: the query is a
    driver call of tens of microseconds)."""
    dev = torch.cuda.current_device()
    if dev not in _DEVICE_BYTES:
        _DEVICE_BYTES[dev] = int(torch.cuda.get_device_properties(dev).total_memory)
    return _DEVICE_BYTES[dev]


def _pool_rays(n, max_children, remaining, gausslet=False):
    """Pool size for a device-resident chain from a generation of n rays:
    room for the generations still allowed (at most _CHAIN_GENERATIONS) at
    about n rays each, plus the worst case the chain has to be sure of before
    it launches a generation without waiting (its own size bound and its
    children's), rounded up to a power of two below 2^24 rays and to a
    multiple of 2^20 above, and capped at 40 % of the device memory free now.
    (A field stride of 2^k rays puts every SoA plane on the same boundary; on
    C2 the generation kernel ran 3-4 % faster with such strides than with
    15 x 2^22 or odd ones.  Large pools round finer, so that the hundreds of
    GB a deep branching trace keeps are not doubled by rounding.)"""
    torch = require_gpu()
    mc = max(int(max_children), 1)
    if remaining <= 1:
        return 1            # the last generation's children are not kept
    want = n * (min(remaining, _CHAIN_GENERATIONS) + 1 + 2 * mc)
    if want <= 1 << 24:
        want = 1 << max(int(want - 1).bit_length(), 0)
    else:
        want = ((want + (1 << 20) - 1) >> 20) << 20
    # memory the pool can come from: the device's memory less what tensors
    # hold now (free or cached by the allocator), so that a repeated trace
    # asks for the same pools as before and gets them from the cache.  The
    # allocator's statistics cost tens of microseconds per query, so they are
    # read only for a pool above a tenth of the device, which the cap could
    # bind (a smaller pool that does not fit is halved by _alloc_pool)
    ray_bytes = _abi.N_F64 * 8 + _abi.N_U32 * 4 + (_abi.N_PARA * _abi.PARA_FIELDS * 8 if gausslet else 0)
    total = _device_bytes(torch)
    if want * ray_bytes <= total // 10:
        return max(want, mc * n, 1)
    cap = int(0.4 * (total - torch.cuda.memory_allocated())) // ray_bytes
    return max(min(want, cap), mc * n, 1)


def _alloc_pool(make, want, minimum):
    """make(capacity) for the largest capacity from `want` down to `minimum`
    (halving) that the device can hold: after an out-of-memory error, the
    allocator's cached blocks are released and a smaller pool is tried (a
    smaller pool only means an earlier restart of the device-resident loop)."""
    torch = require_gpu()
    cap = max(int(want), int(minimum), 1)
    while True:
        try:
            return make(cap)
        except torch.OutOfMemoryError:
            torch.cuda.empty_cache()
            if cap <= minimum:
                raise
            cap = max(cap // 2, int(minimum))


def _trace_chain(ctx, rays, max_length, limit):
    """trace_rays' generation loop for a RayCollection (rpx_trace_rays) or a
    GaussletCollection (rpx_trace_gausslets) on the device: generations are
    written into pools and traced with no host round trip between them; a
    pool that cannot be sure to hold the next generation ends the call, and
    the loop continues from that generation with a fresh pool.  Returns the
    traced collections, the input first."""
    from ._device import DeviceRays, DeviceRaysView
    torch = require_gpu()
    lib = load()
    gauss = isinstance(rays, GaussletCollection)
    cls = GaussletCollection if gauss else RayCollection
    entry = lib.rpx_trace_gausslets if gauss else lib.rpx_trace_rays
    traced = [rays]
    current = rays
    remaining = limit
    hits = (ctypes.c_uint32 * max(ctx.tables.n_global_faces, 1))()
    while remaining > 0 and current.n_rays > 0:
        n = current.n_rays
        dev_in = current._ensure_device()
        wl = _wavelength_table(current, ctx)
        pool = _alloc_pool(lambda cap: DeviceRays(cap, gausslet=gauss),
                           _pool_rays(n, ctx.max_children, remaining, gauss),
                           max(ctx.max_children, 1) * n if remaining > 1 else 1)
        ws_bytes = int(lib.rpx_chain_workspace_bytes(max(n, pool.capacity), remaining))
        ws = ctx.chain_workspace(ws_bytes)
        counts = (ctypes.c_uint64 * (remaining + 1))()
        offsets = (ctypes.c_uint64 * (remaining + 1))()
        n_traced = ctypes.c_int32(0)
        s_in, s_pool = dev_in.struct(), pool.struct()
        code = entry(ctx.handle, ctypes.byref(s_in), n, float(max_length), ctypes.c_void_p(wl.data_ptr()),
                     len(current.wavelengths), remaining, ctypes.byref(s_pool), counts, offsets,
                     ctypes.byref(n_traced), hits, ctypes.c_void_p(ws.data_ptr()), ws_bytes, stream_handle())
        if code != _abi.RPX_E_CAPACITY:
            check(code)
        current._device_changed()
        k_traced = n_traced.value
        # generations 1..k_traced-1 of this call were traced; on a full pool
        # generation k_traced is stored and continues the loop
        last = k_traced + 1 if code == _abi.RPX_E_CAPACITY else k_traced
        parent = current
        for k in range(1, last):
            gen = cls.from_device(DeviceRaysView(pool, offsets[k], counts[k]), counts[k])
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


def trace_rays(input_rays, face_lists, recursion_limit=100, max_length=100.0):
    """Non-sequential trace of a RayCollection or GaussletCollection
    (core/tracer.py:9-47): returns (traced_rays, all_faces).

    As in the reference, each collection is appended before it is traced, so
    with recursion_limit=L the L-th traced generation is not returned (its
    parents still receive their length and end_face_idx)."""
    # the device-resident loop writes every input ray's length itself (the
    # distance to its face, or max_length), so the reset is the kernel's; the
    # gausslet generation reads the input lengths (the reference resets them)
    rays_chain = isinstance(input_rays, RayCollection)
    all_faces = _prepare(input_rays, face_lists, max_length, reset=not rays_chain or input_rays.n_rays == 0
                         or recursion_limit <= 0)
    face_sets = list(face_lists)
    for fs in face_sets:
        fs.sync_transforms()
    ctx = _TraceContext(face_sets, all_faces)
    decomp_faces = [f for f in all_faces if f.material.is_decomp_material()]
    # gausslets that end on a decomposition face are captured and decomposed
    # on the host between generations, so such scenes keep the per-generation
    # loop below
    chain = rays_chain or (isinstance(input_rays, GaussletCollection) and not decomp_faces)
    if chain:
        traced_rays = _trace_chain(ctx, input_rays, max_length, recursion_limit) \
            if input_rays.n_rays > 0 and recursion_limit > 0 else []
        return traced_rays, all_faces
    traced_rays = []
    rays = input_rays
    count = 0
    while rays.n_rays > 0 and count < recursion_limit:
        traced_rays.append(rays)
        rays = _generation(ctx, rays, max_length, flush=False, decomp_faces=decomp_faces)
        count += 1
    ctx.flush_counts()
    return traced_rays, all_faces


def trace_ray_sequence(input_rays, face_sequence, recursion_limit=100, max_length=100.0):
    """Sequential trace through (FaceList, face_idx) pairs (core/tracer.py:50-99)."""
    input_rays.reset_length(max_length)
    traced_rays = [input_rays]
    face_lists = [fl for fl, fidx in face_sequence]
    face_idx_list = [fidx for fl, fidx in face_sequence]
    all_faces = _prepare(input_rays, face_lists, max_length)
    decomp_faces = [f for f in all_faces if f.material.is_decomp_material()]
    contexts = {}
    rays = input_rays
    count = 0
    for fl, fidx in zip(face_lists, face_idx_list):
        fl.sync_transforms()
        key = id(fl)
        if key not in contexts:
            contexts[key] = _TraceContext([fl], all_faces)
        rays = _generation(contexts[key], rays, max_length, 0, fidx, decomp_faces=decomp_faces)
        if (count > recursion_limit) or (rays.n_rays == 0):
            break
        traced_rays.append(rays)
        count += 1
    return traced_rays, all_faces
