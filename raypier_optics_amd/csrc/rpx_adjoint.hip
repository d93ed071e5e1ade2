// Spot RMS of a coaxial sequential lens and its gradient with respect to
// every surface's curvature, conic constant and A4..A10 terms.
//
// Forward (k_lens_forward): one thread per ray traces surfaces 0..n-1 and the
// detector plane and writes per-workgroup moment partials.  Adjoint
// (k_lens_adjoint): each thread traces its ray again, keeping the state
// entering every surface, then runs the chain rule backwards from the
// detector, recomputing each surface's intersection, normal and refraction
// from the stored entering state; surface parameters receive the
// contributions of the implicit intersection equation and of the normal.
// Per-workgroup parameter partials are summed by a fixed-order reduction.
//
// Surface:  G(q) = S(rho) - (q.z - z_s) = 0,  rho = q.x^2 + q.y^2,
//   S(rho)  = c rho / (1 + w) + A4 rho^2 + A6 rho^3 + A8 rho^4 + A10 rho^5,
//   w       = sqrt(1 - (1 + k) c^2 rho),
//   S'(rho) = c / (2 w) + 2 A4 rho + 3 A6 rho^2 + 4 A8 rho^3 + 5 A10 rho^4,
//   normal  = m / |m|,  m = (-2 q.x S', -2 q.y S', 1).
// Refraction (unit d, unit normal N oriented along d, mu = n1 / n2):
//   cos = d.N,  t = mu d + (sqrt(1 - mu^2 (1 - cos^2)) - mu cos) N.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <string>
#include "rpx_internal.h"
#include "rpx_host.h"

using namespace rpx;

namespace {

constexpr int LB = 256;   // workgroup size

struct SurfEval {
    double S, Sp, Spp;          // S, dS/drho, d2S/drho2
    double dS[6], dSp[6];       // d/dparam of S and S'
    bool ok;
};

__device__ __forceinline__ SurfEval surface_eval(const double *P, double rho, bool derivs) {
    SurfEval e;
    const double c = P[0], k = P[1], A4 = P[2], A6 = P[3], A8 = P[4], A10 = P[5];
    const double u = (1 + k) * c * c;
    const double arg = 1 - u * rho;
    e.ok = arg > 0;
    const double w = sqrt(e.ok ? arg : 1.0);
    const double r2 = rho * rho, r3 = r2 * rho, r4 = r3 * rho, r5 = r4 * rho;
    e.S = c * rho / (1 + w) + A4 * r2 + A6 * r3 + A8 * r4 + A10 * r5;
    e.Sp = c / (2 * w) + 2 * A4 * rho + 3 * A6 * r2 + 4 * A8 * r3 + 5 * A10 * r4;
    e.Spp = c * u / (4 * w * w * w) + 2 * A4 + 6 * A6 * rho + 12 * A8 * r2 + 20 * A10 * r3;
    if (derivs) {
        const double w1 = 1 + w;
        e.dS[0] = rho / w1 + u * r2 / (w * w1 * w1);
        e.dS[1] = c * c * c * r2 / (2 * w * w1 * w1);
        e.dS[2] = r2; e.dS[3] = r3; e.dS[4] = r4; e.dS[5] = r5;
        e.dSp[0] = 1 / (2 * w) + u * rho / (2 * w * w * w);
        e.dSp[1] = c * c * c * rho / (4 * w * w * w);
        e.dSp[2] = 2 * rho; e.dSp[3] = 3 * r2; e.dSp[4] = 4 * r3; e.dSp[5] = 5 * r4;
    }
    return e;
}

// Newton on G(t) = S(rho(t)) - (p.z + t d.z - z_s), started at the vertex
// plane.  Converged once a step is below 1e-9 (1 + |t|); two further steps
// then bring the root to float64 rounding (Newton's error squares each step).
__device__ __forceinline__ bool newton_step(const double *P, double zs, V3 p, V3 d, double &t, double &dt) {
    const double qx = p.x + t * d.x, qy = p.y + t * d.y, qz = p.z + t * d.z;
    const double rho = qx * qx + qy * qy;
    const SurfEval e = surface_eval(P, rho, false);
    if (!e.ok) return false;
    const double G = e.S - (qz - zs);
    const double Gt = e.Sp * 2 * (qx * d.x + qy * d.y) - d.z;
    dt = -G / Gt;
    t += dt;
    return true;
}

__device__ __forceinline__ bool intersect(const double *P, double zs, double ap2, V3 p, V3 d, double &t_out) {
    double t = (zs - p.z) / d.z, dt = 0;
    bool conv = false;
    for (int it = 0; it < 60; ++it) {
        if (!newton_step(P, zs, p, d, t, dt)) return false;
        if (fabs(dt) <= 1e-9 * (1 + fabs(t))) {
            conv = true;
            break;
        }
    }
    if (!conv) return false;
    for (int it = 0; it < 2; ++it)
        if (!newton_step(P, zs, p, d, t, dt)) return false;
    if (!(t > 0)) return false;
    const double qx = p.x + t * d.x, qy = p.y + t * d.y;
    if (qx * qx + qy * qy > ap2) return false;
    t_out = t;
    return true;
}

struct Refr {
    V3 N, m;       // unit normal oriented along d, unnormalised gradient normal
    double mlen, sgn, cosv, sr, kappa, mu;
    V3 t;
};

__device__ __forceinline__ bool refract(V3 q, double Sp, V3 d, double mu, Refr &r) {
    r.m = v3(-2 * q.x * Sp, -2 * q.y * Sp, 1.0);
    r.mlen = sqrt(dot(r.m, r.m));
    const V3 N = mul(r.m, 1.0 / r.mlen);
    r.sgn = dot(d, N) >= 0 ? 1.0 : -1.0;
    r.N = mul(N, r.sgn);
    r.cosv = dot(d, r.N);
    const double rr = 1 - mu * mu * (1 - r.cosv * r.cosv);
    if (rr < 0) return false;
    r.sr = sqrt(rr);
    r.kappa = r.sr - mu * r.cosv;
    r.mu = mu;
    r.t = add(mul(d, mu), mul(r.N, r.kappa));
    return true;
}

struct LensArgs {
    rpx_lens_t L;
    rpx_rays_t rays;
    uint64_t n;
    const double *n_table;
    uint64_t n_wl;
};

// Traces one ray; states[s] = (p, d) entering surface s when `keep`.
__device__ __forceinline__ bool trace_ray(const LensArgs &A, uint64_t i, double &px, double &py, V3 *sp, V3 *sd,
                                          bool keep) {
    const double *f = A.rays.f64;
    const uint64_t cap = A.rays.capacity;
    V3 p = v3(f[RPX_OX * cap + i], f[RPX_OY * cap + i], f[RPX_OZ * cap + i]);
    V3 d = norm(v3(f[RPX_DX * cap + i], f[RPX_DY * cap + i], f[RPX_DZ * cap + i]));
    const uint32_t wl = A.rays.u32[RPX_WL_IDX * cap + i];
    if (wl >= A.n_wl) return false;
    const double *nt = A.n_table + (uint64_t)wl * A.L.n_media;
    for (int s = 0; s < A.L.n_surfaces; ++s) {
        if (keep) {
            sp[s] = p;
            sd[s] = d;
        }
        double t;
        if (!intersect(A.L.param[s], A.L.z[s], A.L.aperture[s] * A.L.aperture[s], p, d, t)) return false;
        const V3 q = add(p, mul(d, t));
        const SurfEval e = surface_eval(A.L.param[s], q.x * q.x + q.y * q.y, false);
        Refr r;
        if (!refract(q, e.Sp, d, nt[A.L.medium_before[s]] / nt[A.L.medium_after[s]], r)) return false;
        p = q;
        d = r.t;
    }
    if (!(d.z > 0)) return false;
    const double t = (A.L.detector_z - p.z) / d.z;
    px = p.x + t * d.x;
    py = p.y + t * d.y;
    if (keep) {
        sp[A.L.n_surfaces] = p;
        sd[A.L.n_surfaces] = d;
    }
    return true;
}

// Sum of `v` over the wavefront (fixed butterfly order).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__global__ __launch_bounds__(LB) void k_lens_forward(LensArgs A, double *partials) {
    __shared__ double red[LB / 64][4];
    const uint64_t i = (uint64_t)blockIdx.x * LB + threadIdx.x;
    double m0 = 0, m1 = 0, m2 = 0, m3 = 0;
    if (i < A.n) {
        double px, py;
        if (trace_ray(A, i, px, py, nullptr, nullptr, false)) {
            m0 = 1;
            m1 = px;
            m2 = py;
            m3 = px * px + py * py;
        }
    }
    m0 = wave_sum(m0);
    m1 = wave_sum(m1);
    m2 = wave_sum(m2);
    m3 = wave_sum(m3);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wave][0] = m0;
        red[wave][1] = m1;
        red[wave][2] = m2;
        red[wave][3] = m3;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double v = 0;
        for (int w = 0; w < LB / 64; ++w) v += red[w][threadIdx.x];
        partials[(uint64_t)blockIdx.x * 4 + threadIdx.x] = v;
    }
}

// out[q] = sum over blocks b of partials[b * width + q], fixed order; one
// workgroup per column q.
__global__ __launch_bounds__(256) void k_lens_reduce(const double *partials, uint64_t nb, int width, double *out) {
    __shared__ double red[256];
    const int q = blockIdx.x;
    double v = 0;
    for (uint64_t b = threadIdx.x; b < nb; b += 256) v += partials[b * width + q];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < (unsigned)w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0];
}

__global__ __launch_bounds__(LB) void k_lens_adjoint(LensArgs A, const double *moments, double *partials) {
    __shared__ double gsum[LB / 64][RPX_LENS_MAX_SURFACES * 6];
    const int S = A.L.n_surfaces, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t i = (uint64_t)blockIdx.x * LB + threadIdx.x;
    // spot statistics of the forward pass
    const double cnt = moments[0];
    const double cx = moments[1] / cnt, cy = moments[2] / cnt;
    const double var = moments[3] / cnt - (cx * cx + cy * cy);
    const double sigma = sqrt(var > 0 ? var : 0.0);
    V3 sp[RPX_LENS_MAX_SURFACES + 1], sd[RPX_LENS_MAX_SURFACES + 1];
    double px = 0, py = 0;
    bool ok = (i < A.n) && cnt > 0 && sigma > 0 && trace_ray(A, i, px, py, sp, sd, true);
    // seed: d sigma / d (px, py) = (p - centroid) / (N sigma)
    V3 pb = v3(0, 0, 0), db = v3(0, 0, 0);
    if (ok) {
        const double gx = (px - cx) / (cnt * sigma), gy = (py - cy) / (cnt * sigma);
        const V3 p = sp[S], d = sd[S];
        const double t = (A.L.detector_z - p.z) / d.z;
        pb = v3(gx, gy, 0);
        db = v3(t * gx, t * gy, 0);
        const double tb = gx * d.x + gy * d.y;
        pb.z += -tb / d.z;
        db.z += -tb * t / d.z;
    }
    const double *nt = ok ? A.n_table + (uint64_t)A.rays.u32[RPX_WL_IDX * A.rays.capacity + i] * A.L.n_media
                          : A.n_table;
    for (int s = S - 1; s >= 0; --s) {
        double th[6] = {0, 0, 0, 0, 0, 0};
        if (ok) {
            const double *P = A.L.param[s];
            const V3 p = sp[s], d = sd[s];
            double t = 0;
            intersect(P, A.L.z[s], A.L.aperture[s] * A.L.aperture[s], p, d, t);   // recomputed
            const V3 q = add(p, mul(d, t));
            const double rho = q.x * q.x + q.y * q.y;
            const SurfEval e = surface_eval(P, rho, true);
            Refr r;
            refract(q, e.Sp, d, nt[A.L.medium_before[s]] / nt[A.L.medium_after[s]], r);
            // outputs of the step: q (adjoint pb), r.t (adjoint db)
            const V3 tb = db;
            V3 qb = pb;
            V3 dbar = mul(tb, r.mu);
            V3 Nb = mul(tb, r.kappa);
            const double kb = dot(tb, r.N);
            const double srb = kb;
            double cb = -r.mu * kb;
            const double rrb = srb / (2 * r.sr);
            cb += rrb * 2 * r.mu * r.mu * r.cosv;
            dbar = add(dbar, mul(r.N, cb));
            Nb = add(Nb, mul(d, cb));
            // oriented unit normal -> unnormalised m
            const V3 Nu = mul(r.N, r.sgn), Nub = mul(Nb, r.sgn);
            const V3 mb = mul(sub(Nub, mul(Nu, dot(Nub, Nu))), 1.0 / r.mlen);
            const double Spb = -2 * (q.x * mb.x + q.y * mb.y);
            qb.x += -2 * e.Sp * mb.x;
            qb.y += -2 * e.Sp * mb.y;
            const double rhob = Spb * e.Spp;
            for (int qn = 0; qn < 6; ++qn) th[qn] += Spb * e.dSp[qn];
            qb.x += 2 * q.x * rhob;
            qb.y += 2 * q.y * rhob;
            // intersection q = p + t d with G(t; p, d, param) = 0
            V3 pbar = qb;
            dbar = add(dbar, mul(qb, t));
            const double tbar = dot(qb, d);
            const double Gt = e.Sp * 2 * (q.x * d.x + q.y * d.y) - d.z;
            const double lam = -tbar / Gt;
            const V3 Gp = v3(2 * q.x * e.Sp, 2 * q.y * e.Sp, -1.0);
            pbar = add(pbar, mul(Gp, lam));
            dbar = add(dbar, mul(Gp, lam * t));
            for (int qn = 0; qn < 6; ++qn) th[qn] += lam * e.dS[qn];
            pb = pbar;
            db = dbar;
        }
        for (int qn = 0; qn < 6; ++qn) {
            const double v = wave_sum(th[qn]);
            if (lane == 0) gsum[wave][s * 6 + qn] = v;
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < S * 6; k += LB) {
        double v = 0;
        for (int w = 0; w < LB / 64; ++w) v += gsum[w][k];
        partials[(uint64_t)blockIdx.x * S * 6 + k] = v;
    }
}

rpx_status_t check_lens(const rpx_lens_t *L, const rpx_rays_t *rays, uint64_t n, const double *n_table,
                        uint64_t n_wl) {
    if (!L || !rays) return rpx_set_error(RPX_E_INVALID, "null lens or rays");
    if (L->n_surfaces < 1 || L->n_surfaces > RPX_LENS_MAX_SURFACES)
        return rpx_set_error(RPX_E_INVALID, "lens must have 1..16 surfaces");
    if (L->n_media < 1) return rpx_set_error(RPX_E_INVALID, "lens has no media");
    for (int s = 0; s < L->n_surfaces; ++s)
        if (L->medium_before[s] < 0 || L->medium_before[s] >= L->n_media || L->medium_after[s] < 0 ||
            L->medium_after[s] >= L->n_media)
            return rpx_set_error(RPX_E_INVALID, "surface medium index out of range");
    if (rays->capacity < n || (n && (!rays->f64 || !rays->u32))) return rpx_set_error(RPX_E_INVALID, "rays");
    if (!n_table || n_wl == 0) return rpx_set_error(RPX_E_INVALID, "index table is empty");
    if (n > 0xFFFFFFFFull) return rpx_set_error(RPX_E_UNSUPPORTED, "more than 2^32-1 rays");
    return RPX_OK;
}

}  // namespace

extern "C" uint64_t rpx_lens_workspace_bytes(uint64_t n_rays, int32_t n_surfaces) {
    const uint64_t nb = (n_rays + LB - 1) / LB;
    const uint64_t width = (uint64_t)std::max(4, 6 * std::max(1, (int)n_surfaces));
    return std::max<uint64_t>(nb, 1) * width * 8;
}

extern "C" rpx_status_t rpx_lens_spot(const rpx_lens_t *lens, const rpx_rays_t *rays, uint64_t n_rays,
                                      const double *n_table, uint64_t n_wavelengths, double *moments, void *workspace,
                                      uint64_t workspace_bytes, void *stream) {
    rpx_status_t st = check_lens(lens, rays, n_rays, n_table, n_wavelengths);
    if (st) return st;
    if (!moments) return rpx_set_error(RPX_E_INVALID, "moments is null");
    if (!workspace || workspace_bytes < rpx_lens_workspace_bytes(n_rays, lens->n_surfaces))
        return rpx_set_error(RPX_E_INVALID, "workspace too small");
    const uint64_t nb = std::max<uint64_t>((n_rays + LB - 1) / LB, 1);
    LensArgs A{*lens, *rays, n_rays, n_table, n_wavelengths};
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_lens_forward, dim3((unsigned)nb), dim3(LB), 0, s, A, (double *)workspace);
    RPX_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_lens_reduce, dim3(4), dim3(256), 0, s, (const double *)workspace, nb, 4, moments);
    RPX_HIP_TRY(hipGetLastError());
    return RPX_OK;
}

extern "C" rpx_status_t rpx_lens_spot_grad(const rpx_lens_t *lens, const rpx_rays_t *rays, uint64_t n_rays,
                                           const double *n_table, uint64_t n_wavelengths, const double *moments,
                                           double *grad, void *workspace, uint64_t workspace_bytes, void *stream) {
    rpx_status_t st = check_lens(lens, rays, n_rays, n_table, n_wavelengths);
    if (st) return st;
    if (!moments || !grad) return rpx_set_error(RPX_E_INVALID, "moments or grad is null");
    if (!workspace || workspace_bytes < rpx_lens_workspace_bytes(n_rays, lens->n_surfaces))
        return rpx_set_error(RPX_E_INVALID, "workspace too small");
    const uint64_t nb = std::max<uint64_t>((n_rays + LB - 1) / LB, 1);
    LensArgs A{*lens, *rays, n_rays, n_table, n_wavelengths};
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_lens_adjoint, dim3((unsigned)nb), dim3(LB), 0, s, A, moments, (double *)workspace);
    RPX_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_lens_reduce, dim3(6 * lens->n_surfaces), dim3(256), 0, s, (const double *)workspace, nb,
                       6 * lens->n_surfaces, grad);
    RPX_HIP_TRY(hipGetLastError());
    return RPX_OK;
}
