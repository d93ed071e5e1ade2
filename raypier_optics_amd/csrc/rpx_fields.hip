// Gausslet E-field summation on gfx950.
//
//   k_modes_gc    one gausslet per lane: parabasal offsets/slopes
//                 (evaluate_neighbours_gc, core/fields.py:114-136) fused with
//                 the closed-form 3x3 least-squares mode fit
//                 (evaluate_one_mode, cfields.pyx:156-215).
//   k_field_prep  per-gausslet invariants of sum_gaussian_modes
//                 (cfields.pyx:76-99) into a structure-of-arrays table.
//   k_field_sum   one evaluation point per lane; gausslet invariants stream
//                 through LDS in tiles and every lane sums them in gausslet
//                 order, so each point's sum has the reference's order.
#include <hip/hip_runtime.h>
#include <mutex>
#include <algorithm>
#include <string>
#include "rpx_materials.h"
#include "rpx_host.h"
#include "rpx_gen.h"

using namespace rpx;

// --------------------------------------------------------------- mode fit
RPX_DEV void fit_one_mode(const double *x, const double *y, const double *dx, const double *dy, int size,
                          double blending, double *o) {
    double a00 = 0, a01 = 0, a02 = 0, a11 = 0, a12 = 0, a22 = 0, b0 = 0, b1 = 0, b2 = 0;
    for (int i = 0; i < size; ++i) {
        double xi2 = x[i] * x[i];
        double yi2 = y[i] * y[i];
        a00 += xi2 * xi2;
        a01 += 2 * xi2 * x[i] * y[i];
        a02 += xi2 * yi2;
        a11 += 4 * xi2 * yi2;
        a12 += 2 * x[i] * y[i] * yi2;
        a22 += yi2 * yi2;
        b0 += xi2;
        b1 += 2 * x[i] * y[i];
        b2 += yi2;
    }
    double den = (-a00 * a11 * a22 + a00 * (a12 * a12) + (a01 * a01) * a22 - 2 * a01 * a02 * a12 + (a02 * a02) * a11);
    o[1] = blending * (-b0 * (a11 * a22 - (a12 * a12)) + b1 * (a01 * a22 - a02 * a12) - b2 * (a01 * a12 - a02 * a11)) / den;
    o[3] = -blending * (b0 * (a01 * a22 - a02 * a12) - b1 * (a00 * a22 - (a02 * a02)) + b2 * (a00 * a12 - a01 * a02)) / den;
    o[5] = blending * (-b0 * (a01 * a12 - a02 * a11) + b1 * (a00 * a12 - a01 * a02) - b2 * (a00 * a11 - (a01 * a01))) / den;
    a00 = a01 = a02 = a11 = a12 = a22 = b0 = b1 = b2 = 0;
    for (int i = 0; i < size; ++i) {
        double xi2 = x[i] * x[i];
        double yi2 = y[i] * y[i];
        a00 += xi2;
        a01 += x[i] * y[i];
        a11 += xi2 + yi2;
        a22 += yi2;
        b0 += dx[i] * x[i];
        b1 += dx[i] * y[i] + dy[i] * x[i];
        b2 += dy[i] * y[i];
    }
    a12 = a01 * a01;
    den = (a00 * a12 - a00 * a11 * a22 + a12 * a22);
    o[0] = (-a12 * b2 + a01 * a22 * b1 + b0 * (a12 - a11 * a22)) / den;
    o[2] = -(a00 * a01 * b2 - a00 * a22 * b1 + a01 * a22 * b0) / den;
    o[4] = (a00 * a01 * b1 - a12 * b0 - b2 * (a00 * a11 - a12)) / den;
}

__global__ void k_modes_gc(rpx_rays_t g, uint64_t n, double blending, double *modes) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t c = g.capacity;
    const double *f = g.f64;
    V3 o = v3(f[RPX_OX * c + i], f[RPX_OY * c + i], f[RPX_OZ * c + i]);
    V3 d = v3(f[RPX_DX * c + i], f[RPX_DY * c + i], f[RPX_DZ * c + i]);
    V3 E = v3(f[RPX_EX * c + i], f[RPX_EY * c + i], f[RPX_EZ * c + i]);
    V3 H = cross(E, d);  // numpy.cross(E, direction)
    double x[6], y[6], dx[6], dy[6];
    for (int j = 0; j < RPX_N_PARA; ++j) {
        const uint64_t b = (uint64_t)j * RPX_PARA_FIELDS;
        V3 po = v3(g.para[(b + 0) * c + i], g.para[(b + 1) * c + i], g.para[(b + 2) * c + i]);
        V3 nd = v3(g.para[(b + 3) * c + i], g.para[(b + 4) * c + i], g.para[(b + 5) * c + i]);
        V3 off = sub(po, o);
        double dz = (nd.x * d.x + nd.y * d.y) + nd.z * d.z;
        x[j] = (off.x * E.x + off.y * E.y) + off.z * E.z;
        y[j] = (off.x * H.x + off.y * H.y) + off.z * H.z;
        dx[j] = ((nd.x * E.x + nd.y * E.y) + nd.z * E.z) / dz;
        dy[j] = ((nd.x * H.x + nd.y * H.y) + nd.z * H.z) / dz;
    }
    fit_one_mode(x, y, dx, dy, RPX_N_PARA, blending, modes + 6 * i);
}

#define RPX_MAX_ROW 32

__global__ void k_modes_xy(const double *X, const double *Y, const double *DX, const double *DY, uint64_t n,
                           int row, double blending, double *modes) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fit_one_mode(X + i * row, Y + i * row, DX + i * row, DY + i * row, row, blending, modes + 6 * i);
}

extern "C" rpx_status_t rpx_gausslet_modes(const rpx_rays_t *g, uint64_t n, double blending, double *modes,
                                           void *stream) {
    if (n == 0) return RPX_OK;
    if (!g || !g->f64 || !g->para || !modes || g->capacity < n)
        return rpx_set_error(RPX_E_INVALID, "gausslet buffers");
    hipLaunchKernelGGL(k_modes_gc, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, *g, n,
                       blending, modes);
    RPX_HIP_TRY(hipGetLastError());
    return RPX_OK;
}

extern "C" rpx_status_t rpx_evaluate_modes(const double *x, const double *y, const double *dx, const double *dy,
                                           uint64_t n, int32_t row, double blending, double *modes, void *stream) {
    if (n == 0) return RPX_OK;
    if (!x || !y || !dx || !dy || !modes || row < 1 || row > RPX_MAX_ROW)
        return rpx_set_error(RPX_E_INVALID, "mode-fit buffers or row size");
    hipLaunchKernelGGL(k_modes_xy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, x, y, dx,
                       dy, n, row, blending, modes);
    RPX_HIP_TRY(hipGetLastError());
    return RPX_OK;
}

// ------------------------------------------------------ ray-field modes
// The RayField path of eval_Efield_from_rays (core/fields.py:196-218):
//   k_project_sphere    project_to_sphere (core/fields.py:51-80): rays whose
//                       line meets the sphere move back to the nearer
//                       intersection; the others are dropped (compacted in
//                       ray order).
//   k_ray_field_modes   evaluate_neighbours (core/fields.py:83-111) fused with
//                       the mode fit: rays with six neighbours >= 0 project
//                       each neighbour onto their own origin plane, fit the
//                       mode, and are compacted with it in ray order.
// numpy's sum over a length-3 axis accumulates from +0.0: ((0 + a) + b) + c
// (the leading zero only changes the sign of an exactly zero result).
__device__ __forceinline__ double np_sum3(double a, double b, double c) { return ((0.0 + a) + b) + c; }

__global__ __launch_bounds__(RPX_BLOCK) void k_project_sphere(rpx_rays_t in, uint64_t n, V3 centre, double radius_sq,
                                                              rpx_rays_t out, Workspace w) {
    __shared__ uint32_t lds_wave[RPX_BLOCK / 64];
    __shared__ uint32_t lds_tile;
    __shared__ uint64_t lds_excl;
    if (threadIdx.x == 0) lds_tile = atomicAdd(&w.ctl->tile_counter, 1u);
    __syncthreads();
    const uint32_t tile = lds_tile;
    const uint64_t i = (uint64_t)tile * RPX_BLOCK + threadIdx.x;
    bool sel = false;
    double alpha = 0.0;
    if (i < n) {
        const double *f = in.f64;
        const uint64_t c = in.capacity;
        const V3 o = sub(v3(f[RPX_OX * c + i], f[RPX_OY * c + i], f[RPX_OZ * c + i]), centre);
        const V3 d = v3(f[RPX_DX * c + i], f[RPX_DY * c + i], f[RPX_DZ * c + i]);
        const double cc = np_sum3(o.x * o.x, o.y * o.y, o.z * o.z) - radius_sq;
        const double b = 2 * np_sum3(d.x * o.x, d.y * o.y, d.z * o.z);
        const double disc = b * b - 4 * cc;
        sel = disc >= 0.0;
        if (sel) {
            const double sq = sqrt(disc);
            const double root1 = (-b + sq) / 2, root2 = (-b - sq) / 2;
            alpha = root2 < root1 ? root2 : root1;
        }
    }
    uint32_t block_total;
    const uint32_t local = block_scan_small(sel ? 1u : 0u, lds_wave, &block_total);
    if (threadIdx.x < 64) {
        const uint64_t excl = lookback(w.tile_status, tile, 0, block_total, &w.ctl->error);
        if (threadIdx.x == 0) lds_excl = excl;
    }
    __syncthreads();
    const uint64_t pos = lds_excl + local;
    if (sel && pos < out.capacity) {
        Ray r = load_ray(in, i);
        r.origin = add(r.origin, mul(r.direction, alpha));
        r.accum = r.accum + alpha * r.n.re;
        store_ray(out, pos, r);
    }
    if (threadIdx.x == 0 && (uint64_t)tile == (n - 1) / RPX_BLOCK) w.ctl->children = lds_excl + block_total;
}

__global__ __launch_bounds__(RPX_BLOCK) void k_ray_field_modes(rpx_rays_t in, uint64_t n, const int32_t *nb,
                                                               double blending, rpx_rays_t out, double *modes,
                                                               double *xy, Workspace w) {
    __shared__ uint32_t lds_wave[RPX_BLOCK / 64];
    __shared__ uint32_t lds_tile;
    __shared__ uint64_t lds_excl;
    if (threadIdx.x == 0) lds_tile = atomicAdd(&w.ctl->tile_counter, 1u);
    __syncthreads();
    const uint32_t tile = lds_tile;
    const uint64_t i = (uint64_t)tile * RPX_BLOCK + threadIdx.x;
    bool sel = false;
    double m[6], x[6], y[6], dx[6], dy[6];
    if (i < n) {
        int32_t k6[6];
        sel = true;
        for (int j = 0; j < 6; ++j) {
            k6[j] = nb[i * 6 + j];
            if (k6[j] < 0) sel = false;
        }
        if (sel) {
            for (int j = 0; j < 6; ++j)
                if ((uint64_t)k6[j] >= n) {
                    atomicAdd((unsigned long long *)&w.ctl->hits, 1ull);
                    sel = false;
                }
        }
        if (sel) {
            const double *f = in.f64;
            const uint64_t c = in.capacity;
            const V3 o = v3(f[RPX_OX * c + i], f[RPX_OY * c + i], f[RPX_OZ * c + i]);
            const V3 d = v3(f[RPX_DX * c + i], f[RPX_DY * c + i], f[RPX_DZ * c + i]);
            const V3 E = v3(f[RPX_EX * c + i], f[RPX_EY * c + i], f[RPX_EZ * c + i]);
            const V3 H = cross(E, d);  // numpy.cross(E, direction)
            for (int j = 0; j < 6; ++j) {
                const uint64_t q = (uint64_t)k6[j];
                const V3 no = v3(f[RPX_OX * c + q], f[RPX_OY * c + q], f[RPX_OZ * c + q]);
                const V3 nd = v3(f[RPX_DX * c + q], f[RPX_DY * c + q], f[RPX_DZ * c + q]);
                const V3 off = sub(no, o);
                const double dz = np_sum3(nd.x * d.x, nd.y * d.y, nd.z * d.z);
                const double alpha = (-np_sum3(off.x * d.x, off.y * d.y, off.z * d.z)) / dz;
                const V3 pr = add(off, mul(nd, alpha));
                x[j] = np_sum3(pr.x * E.x, pr.y * E.y, pr.z * E.z);
                y[j] = np_sum3(pr.x * H.x, pr.y * H.y, pr.z * H.z);
                dx[j] = np_sum3(nd.x * E.x, nd.y * E.y, nd.z * E.z) / dz;
                dy[j] = np_sum3(nd.x * H.x, nd.y * H.y, nd.z * H.z) / dz;
            }
            fit_one_mode(x, y, dx, dy, 6, blending, m);
        }
    }
    uint32_t block_total;
    const uint32_t local = block_scan_small(sel ? 1u : 0u, lds_wave, &block_total);
    if (threadIdx.x < 64) {
        const uint64_t excl = lookback(w.tile_status, tile, 0, block_total, &w.ctl->error);
        if (threadIdx.x == 0) lds_excl = excl;
    }
    __syncthreads();
    const uint64_t pos = lds_excl + local;
    if (sel && pos < out.capacity) {
        store_ray(out, pos, load_ray(in, i));
        for (int q = 0; q < 6; ++q) modes[6 * pos + q] = m[q];
        if (xy) {
            for (int q = 0; q < 6; ++q) {
                xy[24 * pos + q] = x[q];
                xy[24 * pos + 6 + q] = y[q];
                xy[24 * pos + 12 + q] = dx[q];
                xy[24 * pos + 18 + q] = dy[q];
            }
        }
    }
    if (threadIdx.x == 0 && (uint64_t)tile == (n - 1) / RPX_BLOCK) w.ctl->children = lds_excl + block_total;
}

static rpx_status_t compact_launch_checks(const rpx_rays_t *in, uint64_t n, const rpx_rays_t *out, uint64_t ws_bytes,
                                          void *ws) {
    if (!in || in->capacity < n) return rpx_set_error(RPX_E_INVALID, "input rays");
    if (n > 0 && (!in->f64 || !in->u32)) return rpx_set_error(RPX_E_INVALID, "input rays have null fields");
    if (!out || (out->capacity > 0 && (!out->f64 || !out->u32))) return rpx_set_error(RPX_E_INVALID, "output rays");
    if (n > 0xFFFFFFFFull) return rpx_set_error(RPX_E_UNSUPPORTED, "more than 2^32-1 rays");
    if (n > 0 && (!ws || ws_bytes < rpx_trace_workspace_bytes(n))) return rpx_set_error(RPX_E_INVALID, "workspace too small");
    return RPX_OK;
}

static rpx_status_t finish_compaction(const Workspace &w, uint64_t capacity, uint64_t *n_out, hipStream_t stream,
                                      const char *what) {
    GenControl ctl;
    RPX_HIP_TRY(read_control(ctl, w.ctl, stream));
    if (ctl.error) return rpx_set_error(RPX_E_HIP, std::string(what) + ": a look-back wait exceeded its poll limit");
    *n_out = ctl.children;
    if (ctl.hits)
        return rpx_set_error(RPX_E_INVALID, std::to_string(ctl.hits) + " rays have a neighbour index outside the rays");
    if (ctl.children > capacity)
        return rpx_set_error(RPX_E_CAPACITY, std::string(what) + ": output capacity below " +
                                                 std::to_string(ctl.children));
    return RPX_OK;
}

extern "C" rpx_status_t rpx_project_to_sphere(const rpx_rays_t *in, uint64_t n, const double *centre, double radius,
                                              rpx_rays_t *out, uint64_t *n_out, void *workspace,
                                              uint64_t workspace_bytes, void *stream_) {
    if (!n_out) return rpx_set_error(RPX_E_INVALID, "n_out is null");
    *n_out = 0;
    if (!centre) return rpx_set_error(RPX_E_INVALID, "centre is null");
    if (rpx_status_t st = compact_launch_checks(in, n, out, workspace_bytes, workspace)) return st;
    if (n == 0) return RPX_OK;
    hipStream_t stream = (hipStream_t)stream_;
    Workspace w = carve(workspace, n);
    const uint64_t nb = (n + RPX_BLOCK - 1) / RPX_BLOCK;
    hipLaunchKernelGGL(k_clear, dim3((unsigned)((std::max<uint64_t>(nb, RPX_SHARDS * RPX_LDS_FACES) + 255) / 256)),
                       dim3(256), 0, stream, w.tile_status, nb, w.ctl);
    RPX_HIP_TRY(hipGetLastError());
    V3 c;
    c.x = centre[0];
    c.y = centre[1];
    c.z = centre[2];
    hipLaunchKernelGGL(k_project_sphere, dim3((unsigned)nb), dim3(RPX_BLOCK), 0, stream, *in, n, c, radius * radius,
                       *out, w);
    RPX_HIP_TRY(hipGetLastError());
    return finish_compaction(w, out->capacity, n_out, stream, "project_to_sphere");
}

extern "C" rpx_status_t rpx_ray_field_modes(const rpx_rays_t *rays, uint64_t n, const int32_t *neighbours,
                                            double blending, rpx_rays_t *out, double *modes, double *neighbour_xy,
                                            uint64_t *n_out, void *workspace, uint64_t workspace_bytes,
                                            void *stream_) {
    if (!n_out) return rpx_set_error(RPX_E_INVALID, "n_out is null");
    *n_out = 0;
    if (rpx_status_t st = compact_launch_checks(rays, n, out, workspace_bytes, workspace)) return st;
    if (n == 0) return RPX_OK;
    if (!neighbours) return rpx_set_error(RPX_E_INVALID, "neighbours are null");
    if (out->capacity > 0 && !modes) return rpx_set_error(RPX_E_INVALID, "modes are null");
    hipStream_t stream = (hipStream_t)stream_;
    Workspace w = carve(workspace, n);
    const uint64_t nb = (n + RPX_BLOCK - 1) / RPX_BLOCK;
    hipLaunchKernelGGL(k_clear, dim3((unsigned)((std::max<uint64_t>(nb, RPX_SHARDS * RPX_LDS_FACES) + 255) / 256)),
                       dim3(256), 0, stream, w.tile_status, nb, w.ctl);
    RPX_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_ray_field_modes, dim3((unsigned)nb), dim3(RPX_BLOCK), 0, stream, *rays, n, neighbours,
                       blending, *out, modes, neighbour_xy, w);
    RPX_HIP_TRY(hipGetLastError());
    return finish_compaction(w, out->capacity, n_out, stream, "ray field modes");
}

// ----------------------------------------------- decomposition mode maths
// calc_mode_curvature (cfields.pyx:231-293): mode basis and curvature
// coefficients from the wavefront's first and second derivatives.
__global__ __launch_bounds__(256) void k_mode_curvature(const double *dx, const double *dy, const double *dx2,
                                                        const double *dy2, const double *dxdy, uint64_t n,
                                                        double *A, double *B, double *C, double *x, double *y,
                                                        double *z) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V3 c = norm(v3(-dx[i], -dy[i], 1.0));
    V3 b = v3(0.0, 0.0, 1.0);
    const V3 a = norm(cross(b, c));
    b = cross(c, a);
    x[3 * i] = a.x; x[3 * i + 1] = a.y; x[3 * i + 2] = a.z;
    y[3 * i] = b.x; y[3 * i + 1] = b.y; y[3 * i + 2] = b.z;
    z[3 * i] = c.x; z[3 * i + 1] = c.y; z[3 * i + 2] = c.z;
    const double Fz = c.x * dx[i] + c.y * dy[i] + c.z;
    A[i] = -(a.x * a.x * dx2[i] + 2 * a.x * a.y * dxdy[i] + a.y * a.y * dy2[i]) / Fz;
    B[i] = -(a.x * b.x * dx2[i] + (a.x * b.y + b.x * a.y) * dxdy[i] + a.y * b.y * dy2[i]) / Fz;
    C[i] = -(b.x * b.x * dx2[i] + 2 * b.x * b.y * dxdy[i] + b.y * b.y * dy2[i]) / Fz;
}

extern "C" rpx_status_t rpx_calc_mode_curvature(const double *dx, const double *dy, const double *dx2,
                                                const double *dy2, const double *dxdy, uint64_t n, double *A,
                                                double *B, double *C, double *x, double *y, double *z,
                                                void *stream) {
    if (n == 0) return RPX_OK;
    if (!dx || !dy || !dx2 || !dy2 || !dxdy || !A || !B || !C || !x || !y || !z)
        return rpx_set_error(RPX_E_INVALID, "mode curvature buffers");
    hipLaunchKernelGGL(k_mode_curvature, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, dx, dy,
                       dx2, dy2, dxdy, n, A, B, C, x, y, z);
    RPX_HIP_TRY(hipGetLastError());
    return RPX_OK;
}

// apply_mode_curvature (cfields.pyx:397-424): tilts each parabasal direction
// by the curvature coefficients in the base ray's (E, E x d) frame.
__global__ __launch_bounds__(256) void k_apply_mode_curvature(rpx_rays_t g, uint64_t n, const double *A,
                                                              const double *B, const double *C) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *f = g.f64;
    const uint64_t c = g.capacity;
    const V3 o = v3(f[RPX_OX * c + i], f[RPX_OY * c + i], f[RPX_OZ * c + i]);
    const V3 zz = v3(f[RPX_DX * c + i], f[RPX_DY * c + i], f[RPX_DZ * c + i]);
    const V3 xx = v3(f[RPX_EX * c + i], f[RPX_EY * c + i], f[RPX_EZ * c + i]);
    const V3 yy = cross(xx, zz);
    const double a = A[i], b = B[i], cc = C[i];
    for (int j = 0; j < RPX_N_PARA; ++j) {
        const uint64_t q = (uint64_t)j * RPX_PARA_FIELDS;
        const V3 po = v3(g.para[(q + 0) * c + i], g.para[(q + 1) * c + i], g.para[(q + 2) * c + i]);
        const V3 pd = v3(g.para[(q + 3) * c + i], g.para[(q + 4) * c + i], g.para[(q + 5) * c + i]);
        const V3 r = sub(po, o);
        const double px = dot(r, xx), py = dot(r, yy);
        V3 shift = mul(xx, (-a * px - b * py));
        shift = add(shift, mul(yy, (-cc * py - b * px)));
        const V3 nd = norm(add(shift, pd));
        g.para[(q + 3) * c + i] = nd.x;
        g.para[(q + 4) * c + i] = nd.y;
        g.para[(q + 5) * c + i] = nd.z;
    }
}

extern "C" rpx_status_t rpx_apply_mode_curvature(rpx_rays_t *gausslets, uint64_t n, const double *A, const double *B,
                                                 const double *C, void *stream) {
    if (n == 0) return RPX_OK;
    if (!gausslets || gausslets->capacity < n || !gausslets->f64 || !gausslets->para || !A || !B || !C)
        return rpx_set_error(RPX_E_INVALID, "mode curvature buffers");
    hipLaunchKernelGGL(k_apply_mode_curvature, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       *gausslets, n, A, B, C);
    RPX_HIP_TRY(hipGetLastError());
    return RPX_OK;
}

// calc_mode_U (cfields.pyx:118-151), statement by statement.
RPX_DEV Cx calc_mode_u(Cx A, Cx B, Cx C, Cx detG0, V3 pt, V3 E, V3 H, V3 dir, Cx k, double phase,
                       double inv_root_area) {
    const Cx I = cx(0.0, 1.0);
    const Cx rootI = cx(sqrt(0.5), sqrt(0.5));
    const double x = dot(pt, E), y = dot(pt, H), z = dot(pt, dir);
    const Cx denom = cmul(cadd(cadd(cx(1, 0), cmul(cx(z, 0), cadd(A, C))), cmul(cx(z * z, 0), detG0)), cx(2, 0));
    const Cx zd = cmul(cx(z, 0), detG0);
    const Cx AA = cdiv(cadd(A, zd), denom);
    const Cx CC = cdiv(cadd(C, zd), denom);
    Cx inner = cadd(cx(z, 0), cmul(AA, cx(x * x, 0)));
    inner = cadd(inner, cdiv(cmul(B, cx(2 * x * y, 0)), denom));
    inner = cadd(inner, cmul(CC, cx(y * y, 0)));
    Cx U = cexp_(cmul(I, cadd(cx(phase, 0), cmul(k, inner))));
    const Cx q1 = cadd(cx(1, 0), cmul(cx(z, 0), A));
    const Cx q2 = cadd(cx(1, 0), cmul(cx(z, 0), C));
    const Cx zb = cmul(cx(z, 0), B);
    U = cdiv(U, csqrt_(cmul(csub(cmul(q1, q2), cmul(zb, zb)), I)));
    return cmul(U, cmul(cx(inv_root_area, 0), rootI));
}

// build_interaction_matrix (cfields.pyx:295-395): the COO entries (i, j,
// E_ij) for every mode i and test point j within max_spacing, in (i, j)
// order -- one lane per pair of the flattened index i * L + j, placed by the
// look-back.
struct InteractionArgs {
    const double *rx, *ry, *tx, *ty, *A, *B, *C, *x, *y, *z;
    uint64_t n, l;
    double k, impart, max2;
    double *data;
    uint64_t *ij;
    uint64_t capacity;
    Workspace ws;
};

__global__ __launch_bounds__(RPX_BLOCK) void k_interaction(InteractionArgs P) {
    __shared__ uint32_t lds_wave[RPX_BLOCK / 64];
    __shared__ uint32_t lds_tile;
    __shared__ uint64_t lds_excl;
    if (threadIdx.x == 0) lds_tile = atomicAdd(&P.ws.ctl->tile_counter, 1u);
    __syncthreads();
    const uint32_t tile = lds_tile;
    const uint64_t total = P.n * P.l;
    const uint64_t t = (uint64_t)tile * RPX_BLOCK + threadIdx.x;
    bool sel = false;
    uint64_t i = 0, j = 0;
    V3 pt = v3(0, 0, 0);
    if (t < total) {
        i = t / P.l;
        j = t - i * P.l;
        pt = sub(v3(P.tx[j], P.ty[j], 0.0), v3(P.rx[i], P.ry[i], 0.0));
        sel = ((pt.x * pt.x) + (pt.y * pt.y)) <= P.max2;
    }
    Cx U = cx(0, 0);
    if (sel) {
        const Cx A = cx(P.A[i], P.impart), B = cx(P.B[i], 0.0), C = cx(P.C[i], P.impart);
        const Cx detG0 = cmul(A, C);
        const V3 a = v3(P.x[3 * i], P.x[3 * i + 1], P.x[3 * i + 2]);
        const V3 b = v3(P.y[3 * i], P.y[3 * i + 1], P.y[3 * i + 2]);
        const V3 c = v3(P.z[3 * i], P.z[3 * i + 1], P.z[3 * i + 2]);
        U = calc_mode_u(A, B, C, detG0, pt, a, b, c, cx(P.k, 0.0), 0.0, 1.0);
    }
    uint32_t block_total;
    const uint32_t local = block_scan_small(sel ? 1u : 0u, lds_wave, &block_total);
    if (threadIdx.x < 64) {
        const uint64_t excl = lookback(P.ws.tile_status, tile, 0, block_total, &P.ws.ctl->error);
        if (threadIdx.x == 0) lds_excl = excl;
    }
    __syncthreads();
    const uint64_t pos = lds_excl + local;
    if (sel && pos < P.capacity) {
        P.data[2 * pos] = U.re;
        P.data[2 * pos + 1] = U.im;
        P.ij[pos] = i;
        P.ij[P.capacity + pos] = j;
    }
    if (threadIdx.x == 0 && (uint64_t)tile == (total - 1) / RPX_BLOCK) P.ws.ctl->children = lds_excl + block_total;
}

extern "C" rpx_status_t rpx_build_interaction_matrix(const double *rx, const double *ry, uint64_t n,
                                                     const double *tx, const double *ty, uint64_t l,
                                                     const double *A, const double *B, const double *C,
                                                     const double *x, const double *y, const double *z,
                                                     double wavelength, double spacing, double max_spacing,
                                                     double blending, double *data, uint64_t *ij,
                                                     uint64_t capacity, uint64_t *n_out, void *workspace,
                                                     uint64_t workspace_bytes, void *stream_) {
    if (!n_out) return rpx_set_error(RPX_E_INVALID, "n_out is null");
    *n_out = 0;
    const uint64_t total = n * l;
    if (total == 0) return RPX_OK;
    if (l != 0 && total / l != n) return rpx_set_error(RPX_E_UNSUPPORTED, "n * l overflows");
    if (total > 0xFFFFFF00ull) return rpx_set_error(RPX_E_UNSUPPORTED, "more than 2^32-256 mode / point pairs");
    if (!rx || !ry || !tx || !ty || !A || !B || !C || !x || !y || !z)
        return rpx_set_error(RPX_E_INVALID, "interaction matrix inputs");
    if (capacity > 0 && (!data || !ij)) return rpx_set_error(RPX_E_INVALID, "interaction matrix outputs");
    if (!workspace || workspace_bytes < rpx_trace_workspace_bytes(total))
        return rpx_set_error(RPX_E_INVALID, "workspace too small");
    hipStream_t stream = (hipStream_t)stream_;
    InteractionArgs P;
    P.rx = rx; P.ry = ry; P.tx = tx; P.ty = ty; P.A = A; P.B = B; P.C = C; P.x = x; P.y = y; P.z = z;
    P.n = n;
    P.l = l;
    P.k = 2000.0 * M_PI / wavelength;
    P.impart = 2 * (blending * blending) / (P.k * spacing * spacing);
    P.max2 = max_spacing * max_spacing;
    P.data = data;
    P.ij = ij;
    P.capacity = capacity;
    P.ws = carve(workspace, total);
    const uint64_t nb = (total + RPX_BLOCK - 1) / RPX_BLOCK;
    hipLaunchKernelGGL(k_clear, dim3((unsigned)((std::max<uint64_t>(nb, RPX_SHARDS * RPX_LDS_FACES) + 255) / 256)),
                       dim3(256), 0, stream, P.ws.tile_status, nb, P.ws.ctl);
    RPX_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_interaction, dim3((unsigned)nb), dim3(RPX_BLOCK), 0, stream, P);
    RPX_HIP_TRY(hipGetLastError());
    return finish_compaction(P.ws, capacity, n_out, stream, "interaction matrix");
}

// ------------------------------------------------------------ field sum FP64
// Per-gausslet invariants, SoA with stride G: origin(3) E(3) H(3) dir(3)
// kz(2) phase A(2) B(2) C(2) detG0(2) inv_root_area E1(2) E2(2) A+C(2)
// V(6) = 36, where V = (E1 E + E2 H) inv_root_area sqrt(i) is the complex
// vector each unit of U adds to the field (cfields.pyx:108-114, 147-148).
#define RPX_GP 36

__global__ void k_field_prep(rpx_rays_t r, uint64_t n, const double *modes, WlTable wl, double ct, double *gp) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Ray ray = load_ray(r, i);
    V3 E = norm(ray.E_vector);
    V3 H = norm(cross(ray.direction, E));
    double k = 2000.0 * M_PI / wl[ray.wl_idx];
    double phase = ray.phase + (ray.accum * k) - (ct * k / ray.n.re);
    Cx kz = cmul(ray.n, cx(k, 0.0));
    double invk = 2. / kz.re;
    Cx A = cx(modes[6 * i], modes[6 * i + 1]);
    Cx B = cx(modes[6 * i + 2], modes[6 * i + 3]);
    Cx C = cx(modes[6 * i + 4], modes[6 * i + 5]);
    double ira = sqrt(sqrt(A.im * C.im - (B.im * B.im)) * (2.0 / M_PI));
    A.im *= invk;
    B.im *= invk;
    C.im *= invk;
    Cx detG0 = csub(cmul(A, C), cmul(B, B));
    const Cx w = cmul(cx(ira, 0.0), cx(sqrt(0.5), sqrt(0.5)));
    const Cx Vx = cmul(cadd(cmul(ray.E1, cx(E.x, 0)), cmul(ray.E2, cx(H.x, 0))), w);
    const Cx Vy = cmul(cadd(cmul(ray.E1, cx(E.y, 0)), cmul(ray.E2, cx(H.y, 0))), w);
    const Cx Vz = cmul(cadd(cmul(ray.E1, cx(E.z, 0)), cmul(ray.E2, cx(H.z, 0))), w);
    double v[RPX_GP] = {ray.origin.x, ray.origin.y, ray.origin.z, E.x, E.y, E.z, H.x, H.y, H.z,
                        ray.direction.x, ray.direction.y, ray.direction.z, kz.re, kz.im, phase,
                        A.re, A.im, B.re, B.im, C.re, C.im, detG0.re, detG0.im, ira,
                        ray.E1.re, ray.E1.im, ray.E2.re, ray.E2.im, A.re + C.re, A.im + C.im,
                        Vx.re, Vx.im, Vy.re, Vy.im, Vz.re, Vz.im};
    for (int k2 = 0; k2 < RPX_GP; ++k2) gp[(uint64_t)k2 * n + i] = v[k2];
}

#define FS_BLOCK 256
#define FS_TILE 64

// Complex product with fused multiply-adds (the field sum's 1e-9 bound
// allows contraction; the trace code in this file keeps separate roundings).
__device__ __forceinline__ Cx cmul_fma(Cx a, Cx b) {
    return cx(fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re));
}

// Rows of the invariants table the float64 pair reads: 0-22 and 28-35.
#define FS_ROWS 31
__device__ __forceinline__ int fs_row(int q) { return q < 23 ? q : q + 5; }

// Block (x, y): points x*FS_BLOCK.., gausslets [y*chunk, min(G, (y+1)*chunk)),
// summed in gausslet order into out + y*6*M.
__global__ __launch_bounds__(FS_BLOCK) void k_field_sum(const double *gp, uint64_t G, uint64_t chunk,
                                                        const double *points, uint64_t M, double *out) {
    __shared__ double tile[FS_ROWS][FS_TILE];
    const uint64_t ip = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t g0 = (uint64_t)blockIdx.y * chunk, g1 = min(G, g0 + chunk);
    out += (uint64_t)blockIdx.y * 6 * M;
    const bool active = ip < M;
    V3 pt0 = active ? v3(points[3 * ip], points[3 * ip + 1], points[3 * ip + 2]) : v3(0, 0, 0);
    double o0r = 0, o0i = 0, o1r = 0, o1i = 0, o2r = 0, o2i = 0;
    for (uint64_t base = g0; base < g1; base += FS_TILE) {
        const int nt = (int)((g1 - base) < FS_TILE ? (g1 - base) : FS_TILE);
        __syncthreads();
        for (int e = threadIdx.x; e < FS_ROWS * FS_TILE; e += FS_BLOCK) {
            const int q = e / FS_TILE, g = e - q * FS_TILE;
            tile[q][g] = (g < nt) ? gp[(uint64_t)fs_row(q) * G + base + g] : 0.0;
        }
        __syncthreads();
        if (!active) continue;
        for (int g = 0; g < nt; ++g) {
            V3 O = v3(tile[0][g], tile[1][g], tile[2][g]);
            V3 E = v3(tile[3][g], tile[4][g], tile[5][g]);
            V3 H = v3(tile[6][g], tile[7][g], tile[8][g]);
            V3 D = v3(tile[9][g], tile[10][g], tile[11][g]);
            Cx kz = cx(tile[12][g], tile[13][g]);
            double phase = tile[14][g];
            Cx A = cx(tile[15][g], tile[16][g]), B = cx(tile[17][g], tile[18][g]);
            Cx C = cx(tile[19][g], tile[20][g]), detG0 = cx(tile[21][g], tile[22][g]);
            const double ApCr = tile[23][g], ApCi = tile[24][g];
            // calc_mode_U (cfields.pyx:118-151), real-by-complex products
            // written out, the three divisions by `denom` as one complex
            // reciprocal, products and sums fused (equal to the reference's
            // expression to a few ulp)
            V3 pt = sub(pt0, O);
            const double x = fma(pt.x, E.x, fma(pt.y, E.y, pt.z * E.z));
            const double y = fma(pt.x, H.x, fma(pt.y, H.y, pt.z * H.z));
            const double z = fma(pt.x, D.x, fma(pt.y, D.y, pt.z * D.z));
            const double z2 = z * z;
            const double dre = 2.0 * fma(z2, detG0.re, fma(z, ApCr, 1.0));
            const double dim = 2.0 * fma(z2, detG0.im, z * ApCi);
            const double dinv = rcp_nr(fma(dre, dre, dim * dim));
            const Cx rd = cx(dre * dinv, -dim * dinv);
            const Cx AA = cmul_fma(cx(fma(z, detG0.re, A.re), fma(z, detG0.im, A.im)), rd);
            const Cx CC = cmul_fma(cx(fma(z, detG0.re, C.re), fma(z, detG0.im, C.im)), rd);
            const Cx Bd = cmul_fma(B, rd);
            const double x2 = x * x, y2 = y * y, xy2 = 2.0 * x * y;
            const double in_re = fma(AA.re, x2, fma(Bd.re, xy2, fma(CC.re, y2, z)));
            const double in_im = fma(AA.im, x2, fma(Bd.im, xy2, CC.im * y2));
            // i * (phase + kz * inner)
            const double ar = phase + fma(kz.re, in_re, -(kz.im * in_im));
            const double ai = fma(kz.re, in_im, kz.im * in_re);
            double sn, cs;
            sincos_cw(ar, &sn, &cs);
            const double ex = exp(-ai);
            // 1 / csqrt(w), w = ((1 + zA)(1 + zC) - (zB)^2) * I: with h = |w|
            // and sq the principal root (Kahan's branches, t = sqrt((|Re w| +
            // h) / 2)), |sq|^2 = h and 1/sq = conj(sq) / h; 1/h and 1/t come
            // from reciprocal square roots, so the pair needs no division
            const Cx q1 = cx(fma(z, A.re, 1.0), z * A.im), q2 = cx(fma(z, C.re, 1.0), z * C.im);
            const Cx zb = cx(z * B.re, z * B.im);
            const Cx w0 = csub(cmul_fma(q1, q2), cmul_fma(zb, zb));
            const double wr = -w0.im, wi = w0.re;
            const double hh = fma(wr, wr, wi * wi);
            const double rh = rsq_nr(hh);                          // 1/h
            const double m = (fabs(wr) + hh * rh) * 0.5;
            const double rt = rsq_nr(m);                           // 1/t
            const double t = m * rt;
            const double qy = wi * rt * 0.5;                       // wi / (2t)
            const Cx sq = wr >= 0.0 ? cx(t, qy) : cx(fabs(qy), copysign(t, wi));
            const double sinv = ex * rh;
            // U = exp(i arg) / sq (inv_root_area sqrt(i) is folded into V)
            const double ur = fma(cs, sq.re, sn * sq.im) * sinv;
            const double ui = fma(sn, sq.re, -(cs * sq.im)) * sinv;
            // out += U V, V = (E1 E + E2 H) inv_root_area sqrt(i)
            const double vxr = tile[25][g], vxi = tile[26][g], vyr = tile[27][g], vyi = tile[28][g];
            const double vzr = tile[29][g], vzi = tile[30][g];
            o0r = fma(ur, vxr, fma(-ui, vxi, o0r));
            o0i = fma(ur, vxi, fma(ui, vxr, o0i));
            o1r = fma(ur, vyr, fma(-ui, vyi, o1r));
            o1i = fma(ur, vyi, fma(ui, vyr, o1i));
            o2r = fma(ur, vzr, fma(-ui, vzi, o2r));
            o2i = fma(ur, vzi, fma(ui, vzr, o2i));
        }
    }
    if (active) {
        out[6 * ip + 0] = o0r; out[6 * ip + 1] = o0i;
        out[6 * ip + 2] = o1r; out[6 * ip + 3] = o1i;
        out[6 * ip + 4] = o2r; out[6 * ip + 5] = o2i;
    }
}

// ------------------------------------------------- field sum FP32-split
// The same pair term with float32 amplitude arithmetic.  What must stay
// exact to a small fraction of a radian is done in float64: the local
// coordinates (x, y, z) of the point in the gausslet's frame (|p - O| is tens
// of mm) and the large phase phi + Re(kz) z (~1e6 rad at 1e5 wavelengths of
// optical path), which is reduced modulo 2 pi in float64 before the float32
// exponential sees it.  The rest -- the Gamma(z) update (one complex
// reciprocal), the quadratic phase terms (tens of rad), exp, sincos and
// 1/sqrt(i w) from hardware reciprocal square roots -- runs in float32,
// with V = (E1 E + E2 H) inv_root_area sqrt(i) from the invariants table;
// contributions are summed in float64.
#define FS32_TILE 128

__global__ __launch_bounds__(FS_BLOCK) void k_field_sum_split(const double *gp, uint64_t G, uint64_t chunk,
                                                              const double *points, uint64_t M, double *out) {
    // float64 rows: origin(3) E(3) H(3) dir(3) kz.re phase; float32 rows:
    // Im kz, A(2) B(2) C(2) detG0(2) A+C(2) V(6)
    __shared__ double t64[14][FS32_TILE];
    __shared__ float t32[17][FS32_TILE];
    const uint64_t ip = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t g0 = (uint64_t)blockIdx.y * chunk, g1 = min(G, g0 + chunk);
    out += (uint64_t)blockIdx.y * 6 * M;
    const bool active = ip < M;
    const double px = active ? points[3 * ip] : 0, py = active ? points[3 * ip + 1] : 0,
                 pz = active ? points[3 * ip + 2] : 0;
    double o0r = 0, o0i = 0, o1r = 0, o1i = 0, o2r = 0, o2i = 0;
    const double TWO_PI = 6.283185307179586;
    // float32 row sources in the float64 invariants table (see k_field_prep)
    const int src32[17] = {13, 15, 16, 17, 18, 19, 20, 21, 22, 28, 29, 30, 31, 32, 33, 34, 35};
    for (uint64_t base = g0; base < g1; base += FS32_TILE) {
        const int nt = (int)((g1 - base) < FS32_TILE ? (g1 - base) : FS32_TILE);
        __syncthreads();
        for (int e = threadIdx.x; e < 17 * FS32_TILE; e += FS_BLOCK) {
            const int q = e / FS32_TILE, g = e - q * FS32_TILE;
            const bool in = g < nt;
            if (q < 14) t64[q][g] = in ? gp[(uint64_t)(q < 12 ? q : (q == 12 ? 12 : 14)) * G + base + g] : 0.0;
            t32[q][g] = in ? (float)gp[(uint64_t)src32[q] * G + base + g] : 0.0f;
        }
        __syncthreads();
        if (!active) continue;
        for (int g = 0; g < nt; ++g) {
            const double dx = px - t64[0][g], dy = py - t64[1][g], dz = pz - t64[2][g];
            const double x = fma(dx, t64[3][g], fma(dy, t64[4][g], dz * t64[5][g]));
            const double y = fma(dx, t64[6][g], fma(dy, t64[7][g], dz * t64[8][g]));
            const double z = fma(dx, t64[9][g], fma(dy, t64[10][g], dz * t64[11][g]));
            const double big = fma(t64[12][g], z, t64[13][g]);        // phase + Re(kz) z
            const double red = fma(-TWO_PI, rint(big * (1.0 / TWO_PI)), big);
            const float zf = (float)z, xf = (float)x, yf = (float)y;
            const float kzi = t32[0][g];
            const float Ar = t32[1][g], Ai = t32[2][g], Br = t32[3][g], Bi = t32[4][g];
            const float Cr = t32[5][g], Ci = t32[6][g], Dr = t32[7][g], Di = t32[8][g];
            // Gamma(z) update with one complex reciprocal of denom
            const float z2 = zf * zf;
            const float dre = 2.0f * fmaf(z2, Dr, fmaf(zf, t32[9][g], 1.0f));
            const float dim = 2.0f * fmaf(z2, Di, zf * t32[10][g]);
            const float dinv = __builtin_amdgcn_rcpf(fmaf(dre, dre, dim * dim));
            const float rr = dre * dinv, ri = -dim * dinv;
            const float aar = fmaf(zf, Dr, Ar), aai = fmaf(zf, Di, Ai);
            const float ccr = fmaf(zf, Dr, Cr), cci = fmaf(zf, Di, Ci);
            const float x2 = xf * xf, y2 = yf * yf, xy2 = 2.0f * xf * yf;
            // quad = (AA x^2 + B 2xy + CC y^2) / denom
            const float nr = fmaf(aar, x2, fmaf(Br, xy2, ccr * y2));
            const float ni = fmaf(aai, x2, fmaf(Bi, xy2, cci * y2));
            const float qr = fmaf(nr, rr, -(ni * ri)), qi = fmaf(nr, ri, ni * rr);
            // arg = phase + kz (z + quad); the part already in `red` is phase + Re(kz) z
            const float kzr = (float)t64[12][g];
            const float arg_re = (float)red + fmaf(kzr, qr, -(kzi * qi));
            const float arg_im = fmaf(kzi, zf, fmaf(kzr, qi, kzi * qr));
            float sn, cs;
            sincos_cwf(arg_re, &sn, &cs);
            const float mag = expf(-arg_im);
            // 1 / sqrt(i w) as conj(sq) / |w| from reciprocal square roots
            const float q1r = fmaf(zf, Ar, 1.0f), q1i = zf * Ai, q2r = fmaf(zf, Cr, 1.0f), q2i = zf * Ci;
            const float zbr = zf * Br, zbi = zf * Bi;
            const float w_re = fmaf(q1r, q2r, -(q1i * q2i)) - fmaf(zbr, zbr, -(zbi * zbi));
            const float w_im = fmaf(q1r, q2i, q1i * q2r) - 2.0f * zbr * zbi;
            const float wr = -w_im, wi = w_re;
            const float hh = fmaf(wr, wr, wi * wi);
            const float rh = __builtin_amdgcn_rsqf(hh);                  // 1/|w|
            const float m = (fabsf(wr) + hh * rh) * 0.5f;
            const float rt = __builtin_amdgcn_rsqf(m);                   // 1/t
            const float t = m * rt, qy = wi * rt * 0.5f;
            const float sqr = wr >= 0.0f ? t : fabsf(qy), sqi = wr >= 0.0f ? qy : copysignf(t, wi);
            const float sinv = mag * rh;
            const float ur = fmaf(cs, sqr, sn * sqi) * sinv, ui = fmaf(sn, sqr, -(cs * sqi)) * sinv;
            // out += U V, V = (E1 E + E2 H) inv_root_area sqrt(i)
            const float vxr = t32[11][g], vxi = t32[12][g], vyr = t32[13][g], vyi = t32[14][g];
            const float vzr = t32[15][g], vzi = t32[16][g];
            o0r += (double)fmaf(ur, vxr, -(ui * vxi));
            o0i += (double)fmaf(ur, vxi, ui * vxr);
            o1r += (double)fmaf(ur, vyr, -(ui * vyi));
            o1i += (double)fmaf(ur, vyi, ui * vyr);
            o2r += (double)fmaf(ur, vzr, -(ui * vzi));
            o2i += (double)fmaf(ur, vzi, ui * vzr);
        }
    }
    if (active) {
        out[6 * ip + 0] = o0r; out[6 * ip + 1] = o0i;
        out[6 * ip + 2] = o1r; out[6 * ip + 3] = o1i;
        out[6 * ip + 4] = o2r; out[6 * ip + 5] = o2i;
    }
}

// out[e] = sum over chunks c (in order) of part[c*n + e].
__global__ void k_field_reduce(const double *part, uint32_t n_chunks, uint64_t n, double *out) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double s = part[e];
    for (uint32_t c = 1; c < n_chunks; ++c) s += part[(uint64_t)c * n + e];
    out[e] = s;
}

// Workgroups of FS_BLOCK threads of `kernel` resident at once on the current
// device (occupancy per CU times CUs), cached per kernel and device.
static uint64_t field_resident(const void *kernel) {
    static std::mutex mu;
    static struct { const void *k; int dev; uint64_t n; } cache[8];
    static int n_cache = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    {
        std::lock_guard<std::mutex> g(mu);
        for (int i = 0; i < n_cache; ++i)
            if (cache[i].k == kernel && cache[i].dev == dev) return cache[i].n;
    }
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, FS_BLOCK, 0) != hipSuccess || per_cu < 1) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) {
        (void)hipGetLastError();
        cus = 1;
    }
    const uint64_t n = (uint64_t)per_cu * (uint64_t)cus;
    std::lock_guard<std::mutex> g(mu);
    if (n_cache < 8) {
        cache[n_cache].k = kernel;
        cache[n_cache].dev = dev;
        cache[n_cache].n = n;
        ++n_cache;
    }
    return n;
}

// Gausslets per chunk for the 2-D field-sum grid: enough chunks that the grid
// is about four rounds of resident workgroups (fewer point blocks than that
// leave CUs idle, and a grid of 1.3 rounds idles most of them in its second
// round), a whole number of tiles per chunk, and a partial-sum buffer of at
// most 512 MB.
static uint64_t field_chunk(const void *kernel, uint64_t G, uint64_t M, uint64_t tile) {
    const uint64_t pb = (M + FS_BLOCK - 1) / FS_BLOCK;
    const uint64_t tiles = (G + tile - 1) / tile;
    uint64_t c = (4 * field_resident(kernel) + pb - 1) / pb;
    const uint64_t cap = (512ull << 20) / (48 * M);
    c = std::max<uint64_t>(1, std::min(c, std::min<uint64_t>(tiles, std::max<uint64_t>(cap, 1))));
    c = std::min<uint64_t>(c, 65535);
    return ((tiles + c - 1) / c) * tile;
}

static rpx_status_t field_sum_checks(const rpx_rays_t *rays, uint64_t n_rays, const double *modes,
                                     const double *points, uint64_t n_pts, double *out, int32_t precision) {
    if (precision != 0 && precision != 1) return rpx_set_error(RPX_E_INVALID, "precision must be 0 or 1");
    if (n_pts == 0) return RPX_OK;
    if (!points || !out) return rpx_set_error(RPX_E_INVALID, "null point/output buffer");
    if (n_rays == 0) return RPX_OK;
    if (!rays || !rays->f64 || !rays->u32 || rays->capacity < n_rays || !modes)
        return rpx_set_error(RPX_E_INVALID, "ray/mode buffers");
    return RPX_OK;
}

// Gausslet chunks on grid y for this precision (see field_chunk).
static uint64_t field_sum_chunk(uint64_t n_rays, uint64_t n_pts, int32_t precision) {
    const void *kern = precision == 0 ? (const void *)k_field_sum : (const void *)k_field_sum_split;
    return field_chunk(kern, n_rays, n_pts, precision == 0 ? FS_TILE : FS32_TILE);
}

// The field-sum launches: invariants into gp (RPX_GP * n_rays doubles), the
// pair kernel over gausslet chunks of `chunk` (partial sums into `part`,
// 6 * n_pts doubles per chunk, when there is more than one chunk), then the
// chunk-ordered reduction into out.  n_rays > 0 and n_pts > 0.
// With `timed`, the pair kernel is bracketed by events whose elapsed time is
// added to its timing slot once the stream has completed.
static hipError_t field_sum_launch(const rpx_rays_t *rays, uint64_t n_rays, const double *modes, WlTable wl,
                                   const double *points, uint64_t n_pts, double time_ps, double *out,
                                   int32_t precision, uint64_t chunk, double *gp, double *part,
                                   hipStream_t stream, bool timed = false) {
    const double ct = 0.299792458 * time_ps;
    hipLaunchKernelGGL(k_field_prep, dim3((unsigned)((n_rays + 255) / 256)), dim3(256), 0, stream, *rays, n_rays,
                       modes, wl, ct, gp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint64_t n_chunks = (n_rays + chunk - 1) / chunk;
    double *dst = n_chunks > 1 ? part : out;
    const dim3 grid((unsigned)((n_pts + FS_BLOCK - 1) / FS_BLOCK), (unsigned)n_chunks);
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (timed && (hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess)) timed = false;
    if (timed) (void)hipEventRecord(ev[0], stream);
    if (precision == 0)
        hipLaunchKernelGGL(k_field_sum, grid, dim3(FS_BLOCK), 0, stream, gp, n_rays, chunk, points, n_pts, dst);
    else
        hipLaunchKernelGGL(k_field_sum_split, grid, dim3(FS_BLOCK), 0, stream, gp, n_rays, chunk, points, n_pts,
                           dst);
    e = hipGetLastError();
    if (timed) {
        float ms = 0;
        if (e == hipSuccess && hipEventRecord(ev[1], stream) == hipSuccess &&
            hipEventSynchronize(ev[1]) == hipSuccess && hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess)
            rpx_timing_add(precision == 0 ? RPX_TIMING_FIELD_SUM : RPX_TIMING_FIELD_SUM_SPLIT, ms,
                           (double)n_rays * (double)n_pts);
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
    }
    if (e != hipSuccess || n_chunks == 1) return e;
    hipLaunchKernelGGL(k_field_reduce, dim3((unsigned)((6 * n_pts + 255) / 256)), dim3(256), 0, stream,
                       (const double *)part, (uint32_t)n_chunks, 6 * n_pts, out);
    return hipGetLastError();
}

extern "C" rpx_status_t rpx_sum_gaussian_modes(const rpx_rays_t *rays, uint64_t n_rays, const double *modes,
                                               const double *wavelengths, uint64_t n_wavelengths,
                                               const double *points, uint64_t n_pts, double time_ps, double *out,
                                               int32_t precision, void *stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    if (rpx_status_t st = field_sum_checks(rays, n_rays, modes, points, n_pts, out, precision)) return st;
    if (n_pts == 0) return RPX_OK;
    if (n_rays == 0) {
        RPX_HIP_TRY(hipMemsetAsync(out, 0, sizeof(double) * 6 * n_pts, stream));
        return RPX_OK;
    }
    WlTable wl;
    wl.wl = wavelengths;
    wl.n = wavelengths ? n_wavelengths : 0;
    const uint64_t chunk = field_sum_chunk(n_rays, n_pts, precision);
    const uint64_t n_chunks = (n_rays + chunk - 1) / chunk;
    double *gp = nullptr, *part = nullptr;
    RPX_HIP_TRY(hipMallocAsync((void **)&gp, sizeof(double) * RPX_GP * n_rays, stream));
    hipError_t e = hipSuccess;
    if (n_chunks > 1) e = hipMallocAsync((void **)&part, sizeof(double) * 6 * n_pts * n_chunks, stream);
    if (e == hipSuccess)
        e = field_sum_launch(rays, n_rays, modes, wl, points, n_pts, time_ps, out, precision, chunk, gp, part, stream,
                             rpx_timing_is_on());
    hipError_t e2 = hipFreeAsync(gp, stream);
    if (part) {
        hipError_t e3 = hipFreeAsync(part, stream);
        if (e2 == hipSuccess) e2 = e3;
    }
    if (e != hipSuccess) return rpx_set_error(RPX_E_HIP, std::string("field sum launch: ") + hipGetErrorString(e));
    if (e2 != hipSuccess) return rpx_set_error(RPX_E_HIP, std::string("field sum free: ") + hipGetErrorString(e2));
    return RPX_OK;
}

// --------------------------------------------- field sum without allocation
// rpx_sum_gaussian_modes_noalloc keeps the invariants table and the chunk
// partial sums in this file-scope device array instead of stream-ordered
// allocations.  Each device holds its own copy until the process ends; host
// code uses the per-device address from hipGetSymbolAddress, never the
// array's name (which is a host-side stand-in).
static __device__ __align__(256) double scratch[RPX_FIELD_SCRATCH_BYTES / sizeof(double)];

#define RPX_MAX_DEVICES 64

// Device address of `scratch` on the current device, looked up once per
// device; nullptr (with the error recorded) if the lookup fails.
static double *field_scratch(rpx_status_t *st) {
    static std::mutex mu;
    static double *ptrs[RPX_MAX_DEVICES] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        *st = rpx_set_error(RPX_E_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
        return nullptr;
    }
    if (dev < 0 || dev >= RPX_MAX_DEVICES) {
        *st = rpx_set_error(RPX_E_INVALID, "device ordinal beyond the no-alloc scratch table");
        return nullptr;
    }
    std::lock_guard<std::mutex> g(mu);
    if (!ptrs[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *st = rpx_set_error(RPX_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return nullptr;
        }
        ptrs[dev] = (double *)p;
    }
    return ptrs[dev];
}

// out[i] = 0 for i < n (a kernel rather than a memset node, so that the call
// can be captured into a graph).
__global__ void k_fill_zero(double *out, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = 0.0;
}

extern "C" uint64_t rpx_field_noalloc_max_rays(void) {
    return RPX_FIELD_SCRATCH_BYTES / (sizeof(double) * RPX_GP);
}

extern "C" rpx_status_t rpx_sum_gaussian_modes_noalloc(const rpx_rays_t *rays, uint64_t n_rays, const double *modes,
                                                       const double *wavelengths, uint64_t n_wavelengths,
                                                       const double *points, uint64_t n_pts, double time_ps,
                                                       double *out, int32_t precision, void *stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    if (rpx_status_t st = field_sum_checks(rays, n_rays, modes, points, n_pts, out, precision)) return st;
    if (n_pts == 0) return RPX_OK;
    if (n_rays > rpx_field_noalloc_max_rays())
        return rpx_set_error(RPX_E_CAPACITY, "rpx_sum_gaussian_modes_noalloc: " + std::to_string(n_rays) +
                                                 " gausslets exceed the scratch limit of " +
                                                 std::to_string(rpx_field_noalloc_max_rays()));
    rpx_status_t st = RPX_OK;
    double *base = field_scratch(&st);
    if (!base) return st;
    if (n_rays == 0) {
        hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((6 * n_pts + 255) / 256)), dim3(256), 0, stream, out,
                           6 * n_pts);
        RPX_HIP_TRY(hipGetLastError());
        return RPX_OK;
    }
    WlTable wl;
    wl.wl = wavelengths;
    wl.n = wavelengths ? n_wavelengths : 0;
    // the invariants table first, the partial sums after it; the chunk count
    // of the allocating entry point when its partial sums fit (then the two
    // entry points sum in the same order), else as many chunks as fit, else
    // one chunk summed straight into out
    const uint64_t gp_doubles = ((RPX_GP * n_rays + 31) / 32) * 32;
    const uint64_t left = RPX_FIELD_SCRATCH_BYTES / sizeof(double) - gp_doubles;
    const uint64_t tile = precision == 0 ? FS_TILE : FS32_TILE;
    uint64_t chunk = field_sum_chunk(n_rays, n_pts, precision);
    uint64_t n_chunks = (n_rays + chunk - 1) / chunk;
    if (n_chunks > 1 && 6 * n_pts * n_chunks > left) {
        const uint64_t fit = left / (6 * n_pts);
        const uint64_t tiles = (n_rays + tile - 1) / tile;
        chunk = fit < 2 ? n_rays : ((tiles + fit - 1) / fit) * tile;
        n_chunks = (n_rays + chunk - 1) / chunk;
    }
    double *part = n_chunks > 1 ? base + gp_doubles : nullptr;
    hipError_t e =
        field_sum_launch(rays, n_rays, modes, wl, points, n_pts, time_ps, out, precision, chunk, base, part, stream);
    if (e != hipSuccess) return rpx_set_error(RPX_E_HIP, std::string("field sum launch: ") + hipGetErrorString(e));
    return RPX_OK;
}
