"""The C-ABI library loads without a GPU and exports every entry point that
include/rpx.h declares; struct layouts agree with the Python bindings."""
import os
import re
import ctypes

import numpy as np

from raypier_optics_amd import _abi, _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "rpx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rpx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    assert "rpx_trace_segment" in names and "rpx_sum_gaussian_modes" in names
    assert names == _lib.EXPORTED


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_struct_sizes_match_bindings():
    lib = _lib.load()   # load() itself cross-checks every struct size
    for name, size in _abi.STRUCT_SIZES.items():
        assert lib.rpx_struct_size(name.encode()) == size
    assert lib.rpx_abi_version() == _abi.ABI_VERSION


def test_ray_record_layouts():
    # ctracer.pxd:40-64 packed records; test_ctracer.py:16-17 checks the same
    assert _abi.ray_dtype.itemsize == 188
    assert _abi.gausslet_dtype.itemsize == 668
    assert [_abi.ray_dtype.fields[f][1] for f in ("E1_amp", "length", "wavelength_idx", "ray_type_id")] == \
        [112, 144, 168, 184]


def test_scene_create_without_device_reports_no_device():
    """Creating a scene needs a GPU; on a GPU-less host the library returns an
    error code instead of crashing."""
    from tests.conftest import gpu_available
    if gpu_available():
        return
    lib = _lib.load()
    d = _abi.SceneDesc()
    h = ctypes.c_void_p()
    code = lib.rpx_scene_create(ctypes.byref(d), ctypes.byref(h))
    assert code in (_abi.RPX_E_NO_DEVICE, _abi.RPX_E_HIP)
    assert b"device" in lib.rpx_last_error().lower() or code == _abi.RPX_E_HIP
