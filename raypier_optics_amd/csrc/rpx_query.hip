// Batched per-object queries: the device side of Face.intersect /
// compute_normal, Shape.point_inside, Distortion.z_offset(_and_gradient) and
// InterfaceMaterial.eval_child_ray of the Python API.
#include <hip/hip_runtime.h>
#include "rpx_materials.h"
#include "rpx_host.h"

using namespace rpx;

__global__ void k_query_face(SceneView S, int rec, int what, const double *p1, const double *p2, uint64_t n,
                             int is_base, double *out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const rpx_face_t &f = S.faces[rec];
    V3 a = v3(p1[3 * i], p1[3 * i + 1], p1[3 * i + 2]);
    if (what == 0) {
        V3 b = v3(p2[3 * i], p2[3 * i + 1], p2[3 * i + 2]);
        out[i] = face_intersect<V_ALL>(S, f, a, b, is_base);
    } else {
        V3 nrm = face_normal<V_ALL>(S, f, a);
        out[3 * i] = nrm.x;
        out[3 * i + 1] = nrm.y;
        out[3 * i + 2] = nrm.z;
    }
}

__global__ void k_query_shape(SceneView S, int root, const double *xy, uint64_t n, int32_t *out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = shape_inside(S, root, xy[2 * i], xy[2 * i + 1]) ? 1 : 0;
}

__global__ void k_query_zernike(SceneView S, int zi, const double *xy, uint64_t n, double *sag, double *grad) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = xy[2 * i], y = xy[2 * i + 1];
    if (sag) sag[i] = zernike_sag(S, zi, x, y);
    if (grad) {
        V3 g = zernike_grad(S, zi, x, y);
        grad[3 * i] = g.x;
        grad[3 * i + 1] = g.y;
        grad[3 * i + 2] = g.z;
    }
}

struct SlotEmit {
    static constexpr int kCap = 3;
    const rpx_rays_t *out;
    uint64_t base;
    int n = 0;
    RPX_DEV void operator()(const Ray &r) {
        if (n < 3) store_ray(*out, base + n, r);
        ++n;
    }
    RPX_DEV void skip() { ++n; }
};

__global__ void k_eval_child(SceneView S, int mat, rpx_rays_t in, uint64_t n, uint32_t idx_base, const double *pts,
                             const double *nrm, const double *tan_, WlTable wl, rpx_rays_t out, uint32_t *counts) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Ray r = load_ray(in, i);
    Orientation o;
    o.normal = v3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
    o.tangent = v3(tan_[3 * i], tan_[3 * i + 1], tan_[3 * i + 2]);
    V3 p = v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    SlotEmit se{&out, 3 * i};
    eval_children<V_ALL>(S.materials[mat], r, idx_base + (uint32_t)i, p, o, wl, se);
    counts[i] = (uint32_t)se.n;
}

static unsigned nblocks(uint64_t n) { return (unsigned)((n + 255) / 256); }

extern "C" rpx_status_t rpx_query_face(const rpx_scene_t *scene, int32_t face_rec, int32_t what, const double *p1,
                                       const double *p2, uint64_t n, int32_t is_base_ray, double *out,
                                       void *stream) {
    if (!scene || face_rec < 0 || face_rec >= scene->view.n_faces)
        return rpx_set_error(RPX_E_INVALID, "face record out of range");
    if (n == 0) return RPX_OK;
    if (!p1 || !out || (what == 0 && !p2)) return rpx_set_error(RPX_E_INVALID, "null query buffer");
    hipLaunchKernelGGL(k_query_face, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, scene->view, face_rec,
                       what, p1, p2, n, is_base_ray, out);
    RPX_HIP_TRY(hipGetLastError());
    return RPX_OK;
}

extern "C" rpx_status_t rpx_query_shape(const rpx_scene_t *scene, int32_t root, const double *xy, uint64_t n,
                                        int32_t *out, void *stream) {
    if (!scene) return rpx_set_error(RPX_E_INVALID, "scene is null");
    {
        rpx_status_t st = rpx_check_shape_program(scene->host_shape_ops.data(), (int)scene->host_shape_ops.size(),
                                                  root);
        if (st != RPX_OK) return st;
    }
    if (n == 0) return RPX_OK;
    if (!xy || !out) return rpx_set_error(RPX_E_INVALID, "null query buffer");
    hipLaunchKernelGGL(k_query_shape, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, scene->view, root, xy, n,
                       out);
    RPX_HIP_TRY(hipGetLastError());
    return RPX_OK;
}

extern "C" rpx_status_t rpx_query_zernike(const rpx_scene_t *scene, int32_t zernike, const double *xy, uint64_t n,
                                          double *sag, double *grad, void *stream) {
    if (!scene || zernike < 0 || zernike >= scene->view.n_zernikes)
        return rpx_set_error(RPX_E_INVALID, "zernike index out of range");
    if (n == 0) return RPX_OK;
    if (!xy) return rpx_set_error(RPX_E_INVALID, "null query buffer");
    hipLaunchKernelGGL(k_query_zernike, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, scene->view, zernike,
                       xy, n, sag, grad);
    RPX_HIP_TRY(hipGetLastError());
    return RPX_OK;
}

extern "C" rpx_status_t rpx_eval_child_rays(const rpx_scene_t *scene, int32_t material, const rpx_rays_t *in,
                                            uint64_t n, uint32_t idx_base, const double *points,
                                            const double *normals, const double *tangents,
                                            const double *wavelengths, uint64_t n_wavelengths, rpx_rays_t *out,
                                            uint32_t *counts, void *stream) {
    if (!scene || material < 0 || material >= scene->view.n_materials)
        return rpx_set_error(RPX_E_INVALID, "material index out of range");
    if (n == 0) return RPX_OK;
    if (!in || !out || !counts || !points || !normals || !tangents)
        return rpx_set_error(RPX_E_INVALID, "null buffer");
    if (in->capacity < n || out->capacity < 3 * n)
        return rpx_set_error(RPX_E_INVALID, "ray buffer capacity");
    WlTable wl;
    wl.wl = wavelengths;
    wl.n = wavelengths ? n_wavelengths : 0;
    hipLaunchKernelGGL(k_eval_child, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, scene->view, material, *in,
                       n, idx_base, points, normals, tangents, wl, *out, counts);
    RPX_HIP_TRY(hipGetLastError());
    return RPX_OK;
}
