"""rpx_sum_gaussian_modes_noalloc (the field sum that allocates nothing,
keeping its invariants table and partial sums in a static device array)
against rpx_sum_gaussian_modes, both called through the C ABI: two calls in a
row, the second smaller, each equal to the allocating entry point's result
bit for bit (same chunking, same summation order).  Also: an empty gausslet
set gives zeros (a fill kernel, no memset), and a set beyond the scratch
limit is refused with RPX_E_CAPACITY before anything runs."""
import ctypes

import numpy as np
import pytest

from raypier_optics_amd import _abi, _lib, workloads
from raypier_optics_amd.ctracer import RayCollection

pytestmark = pytest.mark.gpu


def _gausslets(n, seed):
    rng = np.random.default_rng(seed)
    r = workloads.base_rays(n)
    r["origin"] = np.column_stack([rng.uniform(-0.5, 0.5, (n, 2)), rng.uniform(-0.1, 0.1, n)])
    d = np.column_stack([rng.uniform(-0.05, 0.05, (n, 2)), np.ones(n)])
    r["direction"] = d / np.linalg.norm(d, axis=1)[:, None]
    e = np.cross(r["direction"], (0.0, 1.0, 0.0))
    r["E_vector"] = e / np.linalg.norm(e, axis=1)[:, None]
    r["E1_amp"] = rng.normal(size=n) + 1j * rng.normal(size=n)
    r["E2_amp"] = rng.normal(size=n) + 1j * rng.normal(size=n)
    r["refractive_index"] = 1.0
    r["accumulated_path"] = rng.uniform(0.0, 100.0, n)
    r["phase"] = rng.uniform(-3, 3, n)
    r["wavelength_idx"] = rng.integers(0, 2, n)
    modes = np.column_stack([rng.uniform(-20, 20, n) + 1j * rng.uniform(800, 1200, n),
                             rng.uniform(-5, 5, n) + 1j * rng.uniform(-50, 50, n),
                             rng.uniform(-20, 20, n) + 1j * rng.uniform(800, 1200, n)])
    return RayCollection.from_array(r), modes


def _call(fn, rays, modes, wl, pts, precision, torch):
    dev = rays._ensure_device()
    s = dev.struct()
    md = torch.from_numpy(np.ascontiguousarray(modes).view(np.float64).copy()).to("cuda")
    w = torch.tensor(wl, dtype=torch.float64, device="cuda")
    p = torch.from_numpy(np.ascontiguousarray(pts, np.float64)).to("cuda")
    npt = len(pts)
    out = torch.full((npt, 3, 2), np.nan, dtype=torch.float64, device="cuda")
    code = fn(ctypes.byref(s), rays.n_rays, ctypes.c_void_p(md.data_ptr()), ctypes.c_void_p(w.data_ptr()),
              w.numel(), ctypes.c_void_p(p.data_ptr()), npt, 0.25, ctypes.c_void_p(out.data_ptr()), precision,
              _lib.stream_handle())
    _lib.check(code)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_matches_allocating_entry(gpu, precision):
    torch = gpu
    lib = _lib.load()
    wl = [1.064, 0.9]
    rng = np.random.default_rng(11)
    for step, (n_g, n_pts) in enumerate(((20000, 4096), (3001, 777))):
        print("no-alloc field sum: precision %d call %d, %d gausslets x %d points" % (precision, step, n_g, n_pts),
              flush=True)
        rays, modes = _gausslets(n_g, 100 + step)
        pts = np.column_stack([rng.uniform(-0.5, 0.5, (n_pts, 2)), rng.uniform(0.5, 3.0, n_pts)])
        want = _call(lib.rpx_sum_gaussian_modes, rays, modes, wl, pts, precision, torch)
        got = _call(lib.rpx_sum_gaussian_modes_noalloc, rays, modes, wl, pts, precision, torch)
        assert np.isfinite(want).all()
        assert np.array_equal(got, want)


def test_noalloc_empty_and_limit(gpu):
    torch = gpu
    lib = _lib.load()
    rays, modes = _gausslets(8, 3)
    pts = np.zeros((100, 3))
    rays_empty = RayCollection.from_array(workloads.base_rays(0))
    got = _call(lib.rpx_sum_gaussian_modes_noalloc, rays_empty, modes[:0], [1.0], pts, 0, torch)
    assert np.array_equal(got, np.zeros_like(got))
    limit = int(lib.rpx_field_noalloc_max_rays())
    assert limit * 36 * 8 <= _abi.RPX_FIELD_SCRATCH_BYTES
    s = rays._ensure_device().struct()
    big = _abi.Rays(s.f64, s.u32, s.para, limit + 1)
    dummy = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = ctypes.c_void_p(dummy.data_ptr())
    code = lib.rpx_sum_gaussian_modes_noalloc(ctypes.byref(big), limit + 1, p, p, 1, p, 1, 0.0, p, 0,
                                              _lib.stream_handle())
    assert code == _abi.RPX_E_CAPACITY
    assert b"scratch" in lib.rpx_last_error()
