"""Per-object GPU queries backing Face.intersect / compute_normal,
Shape.point_inside, ZernikeDistortion.z_offset(_and_gradient) and
InterfaceMaterial.eval_child_ray of the Python API."""
import ctypes

import numpy as np

from . import _abi
from ._lib import load, check, require_gpu, stream_handle
from ._scene import SceneBuilder


def _dev(a, dtype=np.float64):
    torch = require_gpu()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to("cuda")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def face_scene(face):
    from .ctracer import FaceList
    b = SceneBuilder()
    fl = FaceList()
    fl.faces = [face]
    b.add_face_list(fl)
    return b.finish()


def face_intersect(face, p1, p2, is_base_ray):
    """Face.intersect_c distance for each (p1, p2) pair (local frame)."""
    t = face_scene(face)
    torch = require_gpu()
    a, b = _dev(np.reshape(p1, (-1, 3))), _dev(np.reshape(p2, (-1, 3)))
    n = a.shape[0]
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    check(load().rpx_query_face(t.handle(), 0, 0, _ptr(a), _ptr(b), n, int(is_base_ray), _ptr(out),
                                stream_handle()))
    return out.cpu().numpy()


def face_normal(face, p):
    """Face.compute_normal_c at each local point."""
    t = face_scene(face)
    torch = require_gpu()
    a = _dev(np.reshape(p, (-1, 3)))
    n = a.shape[0]
    out = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    check(load().rpx_query_face(t.handle(), 0, 1, _ptr(a), None, n, 1, _ptr(out), stream_handle()))
    return out.cpu().numpy()


def shape_inside(shape, x, y):
    b = SceneBuilder()
    root = b.shape_root(shape)
    t = b.finish()
    torch = require_gpu()
    xy = _dev(np.stack([np.ravel(x), np.ravel(y)], axis=1))
    n = xy.shape[0]
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    check(load().rpx_query_shape(t.handle(), root, _ptr(xy), n, _ptr(out), stream_handle()))
    return out.cpu().numpy().astype(bool)


def zernike_eval(dist, x, y):
    """(sag[n], grad[n, 3]) of a distortion at the points (x, y)."""
    b = SceneBuilder()
    zi = b.distortion_index(dist)
    t = b.finish()
    torch = require_gpu()
    xy = _dev(np.stack([np.ravel(x), np.ravel(y)], axis=1))
    n = xy.shape[0]
    sag = torch.empty(n, dtype=torch.float64, device="cuda")
    grad = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    check(load().rpx_query_zernike(t.handle(), zi, _ptr(xy), n, _ptr(sag), _ptr(grad), stream_handle()))
    return sag.cpu().numpy(), grad.cpu().numpy()


def eval_child_rays_batch(material, rays_aos, idx_base, points, normals, tangents, wavelengths=None):
    """Children of many independent incident rays: returns (children ray_dtype
    array in input order, counts per input)."""
    from ._device import DeviceRays
    b = SceneBuilder()
    mi = b.material_index(material)
    t = b.finish()
    torch = require_gpu()
    n = len(rays_aos)
    din = DeviceRays(max(n, 1))
    din.upload(rays_aos)
    dout = DeviceRays(max(3 * n, 1))
    wl = material._wavelengths if wavelengths is None else np.asarray(wavelengths, np.float64)
    wld = _dev(wl if len(wl) else np.zeros(1))
    pts, nrm, tan = _dev(np.reshape(points, (-1, 3))), _dev(np.reshape(normals, (-1, 3))), \
        _dev(np.reshape(tangents, (-1, 3)))
    counts = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    s_in, s_out = din.struct(), dout.struct()
    check(load().rpx_eval_child_rays(t.handle(), mi, ctypes.byref(s_in), n, int(idx_base), _ptr(pts), _ptr(nrm),
                                     _ptr(tan), _ptr(wld), len(wl), ctypes.byref(s_out), _ptr(counts),
                                     stream_handle()))
    c = counts.cpu().numpy()[:n].astype(np.int64)
    allrec = dout.download(3 * n)
    sel = [3 * i + k for i in range(n) for k in range(min(int(c[i]), 3))]
    return allrec[sel], c


def eval_child_rays(material, ray, idx, point, normal, tangent):
    from .ctracer import Ray
    recs, _ = eval_child_rays_batch(material, ray._a, idx, [point], [normal], [tangent])
    out = []
    for r in recs:
        x = Ray()
        x._a[0] = r
        out.append(x)
    return out
