"""Device storage of ray generations: structure-of-arrays in HBM.

One generation of rays is a single uint8 allocation holding 21 float64 rows
then 5 uint32 rows of `capacity` elements each (see rpx_rays_t); gausslets add
a float64 block of 60 rows (6 parabasal rays x 10 doubles).
"""
import ctypes
import numpy as np

from . import _abi
from ._lib import load, check, require_gpu, stream_handle


class DeviceRays(object):
    """HBM-resident SoA buffer for up to `capacity` rays (or gausslets)."""

    def __init__(self, capacity, gausslet=False):
        torch = require_gpu()
        self.capacity = int(capacity)
        self.gausslet = bool(gausslet)
        nbytes = self.capacity * (_abi.N_F64 * 8 + _abi.N_U32 * 4)
        self.buf = torch.empty(max(nbytes, 8), dtype=torch.uint8, device="cuda")
        if gausslet:
            self.para = torch.empty(max(self.capacity * _abi.N_PARA * _abi.PARA_FIELDS, 1),
                                    dtype=torch.float64, device="cuda")
        else:
            self.para = None

    # torch views -------------------------------------------------------
    def f64(self):
        cap = self.capacity
        return self.buf[: _abi.N_F64 * 8 * cap].view(dtype=_torch().float64).view(_abi.N_F64, cap)

    def u32(self):
        cap = self.capacity
        start = _abi.N_F64 * 8 * cap
        return self.buf[start: start + _abi.N_U32 * 4 * cap].view(dtype=_torch().int32).view(_abi.N_U32, cap)

    def para_view(self):
        cap = self.capacity
        return self.para[: _abi.N_PARA * _abi.PARA_FIELDS * cap].view(_abi.N_PARA, _abi.PARA_FIELDS, cap)

    def struct(self):
        base = self.buf.data_ptr()
        return _abi.Rays(base, base + _abi.N_F64 * 8 * self.capacity,
                         self.para.data_ptr() if self.para is not None else 0, self.capacity)

    # AoS conversion ----------------------------------------------------
    def upload(self, aos):
        """Copy a host ray_dtype / gausslet_dtype array into the SoA buffer."""
        torch = _torch()
        n = len(aos)
        if n == 0:
            return
        raw = np.ascontiguousarray(aos).view(np.uint8).reshape(-1)
        dev = torch.from_numpy(raw.copy()).to("cuda")
        lib = load()
        s = self.struct()
        check(lib.rpx_unpack_rays(ctypes.c_void_p(dev.data_ptr()), n, ctypes.byref(s), int(self.gausslet),
                                  stream_handle()))
        torch.cuda.current_stream().synchronize()

    def download(self, n):
        """Return the first n records as a host ray_dtype / gausslet_dtype array."""
        torch = _torch()
        dt = _abi.gausslet_dtype if self.gausslet else _abi.ray_dtype
        if n == 0:
            return np.zeros(0, dt)
        aos = torch.empty(n * dt.itemsize, dtype=torch.uint8, device="cuda")
        lib = load()
        s = self.struct()
        check(lib.rpx_pack_rays(ctypes.byref(s), n, ctypes.c_void_p(aos.data_ptr()), int(self.gausslet),
                                stream_handle()))
        host = aos.cpu().numpy()
        return host.view(dt).copy()


class DeviceRaysView(DeviceRays):
    """Rays [offset, offset + n) of a pool DeviceRays: the same buffer and
    field stride, so a generation written into a pool is a collection of its
    own without a copy.  Holds a reference to the pool."""

    def __init__(self, pool, offset, n):
        self.pool = pool
        self.offset = int(offset)
        self.capacity = int(n)
        self.gausslet = pool.gausslet
        self.buf = pool.buf
        self.para = pool.para

    def f64(self):
        return self.pool.f64()[:, self.offset: self.offset + self.capacity]

    def u32(self):
        return self.pool.u32()[:, self.offset: self.offset + self.capacity]

    def para_view(self):
        return self.pool.para_view()[:, :, self.offset: self.offset + self.capacity]

    def struct(self):
        s = self.pool.struct()
        o = self.offset
        return _abi.Rays(s.f64 + 8 * o, s.u32 + 4 * o, (s.para + 8 * o) if s.para else 0, self.pool.capacity)


def _torch():
    import torch
    return torch
