// Spot RMS or wavefront RMS of a coaxial sequential lens and its gradient
// with respect to every surface's curvature, conic constant, A4..A10 terms
// and Zernike distortion coefficients.
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
// Surface:  G(q) = S(rho) + D(q.x, q.y) - (q.z - z_s) = 0,
//   rho     = q.x^2 + q.y^2,
//   S(rho)  = c rho / (1 + w) + A4 rho^2 + A6 rho^3 + A8 rho^4 + A10 rho^5,
//   w       = sqrt(1 - (1 + k) c^2 rho),
//   S'(rho) = c / (2 w) + 2 A4 rho + 3 A6 rho^2 + 4 A8 rho^3 + 5 A10 rho^4,
//   D       = sum_k C_k (x/Ru)^a_k (y/Ru)^b_k   (0 without a distortion),
//   normal  = m / |m|,  m = (-(2 q.x S' + D_x), -(2 q.y S' + D_y), 1).
// Wavefront: W = accumulated_path + sum_s n_s t_s + n_L (P - q_L) . d_L.
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

// Monomial k = d(d+1)/2 + (d - a) is x^a y^b with b = d - a, d = 0..6.
constexpr int ZDEG = 6;

// D and its first and second derivatives in x, y (physical units) from the
// monomial coefficients C (normalised coordinates x/Ru, y/Ru).
struct ZEval {
    double D, Dx, Dy, Dxx, Dxy, Dyy;
};

__device__ __forceinline__ void zpowers(double x, double *p) {
    p[0] = 1.0;
#pragma unroll
    for (int i = 1; i <= ZDEG; ++i) p[i] = p[i - 1] * x;
}

__device__ __forceinline__ ZEval zernike_eval(const double *C, double Ru, double x, double y, bool second) {
    ZEval z = {0, 0, 0, 0, 0, 0};
    if (!C) return z;
    const double inv = 1.0 / Ru;
    double xp[ZDEG + 1], yp[ZDEG + 1];
    zpowers(x * inv, xp);
    zpowers(y * inv, yp);
    int k = 0;
#pragma unroll
    for (int d = 0; d <= ZDEG; ++d) {
#pragma unroll
        for (int a = d; a >= 0; --a, ++k) {
            const int b = d - a;
            const double c = C[k];
            z.D += c * xp[a] * yp[b];
            if (a >= 1) z.Dx += c * a * xp[a - 1] * yp[b];
            if (b >= 1) z.Dy += c * b * xp[a] * yp[b - 1];
            if (second) {
                if (a >= 2) z.Dxx += c * a * (a - 1) * xp[a - 2] * yp[b];
                if (a >= 1 && b >= 1) z.Dxy += c * a * b * xp[a - 1] * yp[b - 1];
                if (b >= 2) z.Dyy += c * b * (b - 1) * xp[a] * yp[b - 2];
            }
        }
    }
    z.Dx *= inv;
    z.Dy *= inv;
    z.Dxx *= inv * inv;
    z.Dxy *= inv * inv;
    z.Dyy *= inv * inv;
    return z;
}

// Row of surface s of the monomial table, or null without a distortion.
__device__ __forceinline__ const double *zrow(const rpx_lens_t &L, const double *zc, int s) {
    return (zc && L.zernike_radius[s] > 0) ? zc + (uint64_t)s * RPX_LENS_ZTERMS : nullptr;
}

// Newton on G(t) = S(rho(t)) + D(t) - (p.z + t d.z - z_s), started at the
// vertex plane.  Converged once a step is below 1e-9 (1 + |t|); two further steps
// then bring the root to float64 rounding (Newton's error squares each step).
__device__ __forceinline__ bool newton_step(const double *P, const double *C, double Ru, double zs, V3 p, V3 d,
                                            double &t, double &dt) {
    const double qx = p.x + t * d.x, qy = p.y + t * d.y, qz = p.z + t * d.z;
    const double rho = qx * qx + qy * qy;
    const SurfEval e = surface_eval(P, rho, false);
    if (!e.ok) return false;
    const ZEval z = zernike_eval(C, Ru, qx, qy, false);
    const double G = e.S + z.D - (qz - zs);
    const double Gt = (2 * qx * e.Sp + z.Dx) * d.x + (2 * qy * e.Sp + z.Dy) * d.y - d.z;
    dt = -G / Gt;
    t += dt;
    return true;
}

__device__ __forceinline__ bool intersect(const double *P, const double *C, double Ru, double zs, double ap2, V3 p,
                                          V3 d, double &t_out) {
    double t = (zs - p.z) / d.z, dt = 0;
    bool conv = false;
    for (int it = 0; it < 60; ++it) {
        if (!newton_step(P, C, Ru, zs, p, d, t, dt)) return false;
        if (fabs(dt) <= 1e-9 * (1 + fabs(t))) {
            conv = true;
            break;
        }
    }
    if (!conv) return false;
    for (int it = 0; it < 2; ++it)
        if (!newton_step(P, C, Ru, zs, p, d, t, dt)) return false;
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

__device__ __forceinline__ bool refract(V3 q, double Sp, const ZEval &z, V3 d, double mu, Refr &r) {
    r.m = v3(-(2 * q.x * Sp + z.Dx), -(2 * q.y * Sp + z.Dy), 1.0);
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
    const double *zc;    // monomial table (n_surfaces x RPX_LENS_ZTERMS) or null
    int objective;
};

// Traces one ray: detector point (px, py), its wavefront path W and
// wavelength index; states[s] = (p, d) entering surface s when `keep`.
template <bool ZK, bool WF>
__device__ __forceinline__ bool trace_ray(const LensArgs &A, uint64_t i, double &px, double &py, double &W,
                                          uint32_t &wl, V3 *sp, V3 *sd, bool keep) {
    const double *f = A.rays.f64;
    const uint64_t cap = A.rays.capacity;
    V3 p = v3(f[RPX_OX * cap + i], f[RPX_OY * cap + i], f[RPX_OZ * cap + i]);
    V3 d = norm(v3(f[RPX_DX * cap + i], f[RPX_DY * cap + i], f[RPX_DZ * cap + i]));
    wl = A.rays.u32[RPX_WL_IDX * cap + i];
    if (wl >= A.n_wl) return false;
    const double *nt = A.n_table + (uint64_t)wl * A.L.n_media;
    if constexpr (WF) W = f[RPX_ACCUM * cap + i];
    for (int s = 0; s < A.L.n_surfaces; ++s) {
        if (keep) {
            sp[s] = p;
            sd[s] = d;
        }
        const double *C = ZK ? zrow(A.L, A.zc, s) : nullptr;
        double t;
        if (!intersect(A.L.param[s], C, A.L.zernike_radius[s], A.L.z[s], A.L.aperture[s] * A.L.aperture[s], p, d,
                       t))
            return false;
        const V3 q = add(p, mul(d, t));
        const SurfEval e = surface_eval(A.L.param[s], q.x * q.x + q.y * q.y, false);
        const ZEval z = zernike_eval(C, A.L.zernike_radius[s], q.x, q.y, false);
        Refr r;
        if (!refract(q, e.Sp, z, d, nt[A.L.medium_before[s]] / nt[A.L.medium_after[s]], r)) return false;
        if constexpr (WF) W += nt[A.L.medium_before[s]] * t;
        p = q;
        d = r.t;
    }
    if (!(d.z > 0)) return false;
    const double t = (A.L.detector_z - p.z) / d.z;
    px = p.x + t * d.x;
    py = p.y + t * d.y;
    if constexpr (WF) {
        const double nL = nt[A.L.medium_after[A.L.n_surfaces - 1]];
        W += nL * (-p.x * d.x - p.y * d.y + (A.L.detector_z - p.z) * d.z);
    }
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

// Per-workgroup partials of the objective's moments (ZK: some surface has a
// distortion).  Spot RMS (width 4):
// kept rays, sum x, sum y, sum x^2 + y^2.  Wavefront RMS, pass 0 (width
// 2 n_wl): per wavelength, kept rays and sum W; pass 1 (width 2): kept rays
// and sum (W - mean_l)^2 with the per-wavelength means of pass 0 (read from
// moments).  Sums within a workgroup run in a fixed order.
template <bool ZK, bool WF>
__global__ __launch_bounds__(LB) void k_lens_forward(LensArgs A, int pass, const double *moments, double *partials) {
    __shared__ double red[LB / 64][2 * RPX_LENS_MAX_WFE_WAVELENGTHS];
    const uint64_t i = (uint64_t)blockIdx.x * LB + threadIdx.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double px = 0, py = 0, W = 0;
    uint32_t wl = 0;
    const bool ok = i < A.n && trace_ray<ZK, WF>(A, i, px, py, W, wl, nullptr, nullptr, false);
    int width;
    if constexpr (!WF) {
        width = 4;
        const double v[4] = {ok ? 1.0 : 0.0, ok ? px : 0.0, ok ? py : 0.0, ok ? px * px + py * py : 0.0};
        for (int q = 0; q < 4; ++q) {
            const double t = wave_sum(v[q]);
            if (lane == 0) red[wave][q] = t;
        }
    } else if (pass == 0) {
        width = 2 * (int)A.n_wl;
        for (int l = 0; l < (int)A.n_wl; ++l) {
            const bool m = ok && wl == (uint32_t)l;
            const double c = wave_sum(m ? 1.0 : 0.0), w = wave_sum(m ? W : 0.0);
            if (lane == 0) {
                red[wave][2 * l] = c;
                red[wave][2 * l + 1] = w;
            }
        }
    } else {
        width = 2;
        double dv = 0;
        if (ok) {
            const double mu = moments[3 + 2 * wl] / moments[2 + 2 * wl];
            dv = (W - mu) * (W - mu);
        }
        const double c = wave_sum(ok ? 1.0 : 0.0), v = wave_sum(dv);
        if (lane == 0) {
            red[wave][0] = c;
            red[wave][1] = v;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < width) {
        double v = 0;
        for (int w = 0; w < LB / 64; ++w) v += red[w][threadIdx.x];
        partials[(uint64_t)blockIdx.x * width + threadIdx.x] = v;
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

// Objective value from its moments.
__device__ __forceinline__ double objective_value(int objective, const double *m) {
    if (objective == RPX_LENS_SPOT_RMS) {
        const double cx = m[1] / m[0], cy = m[2] / m[0];
        const double var = m[3] / m[0] - (cx * cx + cy * cy);
        return sqrt(var > 0 ? var : 0.0);
    }
    const double var = m[1] / m[0];
    return sqrt(var > 0 ? var : 0.0);
}

// Partials per workgroup: width n_surfaces * (6 + RPX_LENS_ZTERMS), the
// parameter block then the monomial block.
template <bool ZK, bool WF>
__device__ __forceinline__ void lens_adjoint(const LensArgs &A, const double *moments, double *partials) {
    __shared__ double gsum[LB / 64][RPX_LENS_MAX_SURFACES * (6 + RPX_LENS_ZTERMS)];
    const int S = A.L.n_surfaces, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int zbase = 6 * S;
    const uint64_t i = (uint64_t)blockIdx.x * LB + threadIdx.x;
    const double cnt = moments[0];
    const double obj = cnt > 0 ? objective_value(WF ? RPX_LENS_WAVEFRONT_RMS : RPX_LENS_SPOT_RMS, moments) : 0.0;
    V3 sp[RPX_LENS_MAX_SURFACES + 1], sd[RPX_LENS_MAX_SURFACES + 1];
    double px = 0, py = 0, W = 0;
    uint32_t wl = 0;
    bool ok = (i < A.n) && cnt > 0 && obj > 0 && trace_ray<ZK, WF>(A, i, px, py, W, wl, sp, sd, true);
    const double *nt = ok ? A.n_table + (uint64_t)wl * A.L.n_media : A.n_table;
    // seeds: adjoints of (p, d) leaving the last surface, and the weight of
    // each segment's optical path (wavefront objective)
    V3 pb = v3(0, 0, 0), db = v3(0, 0, 0);
    double omega = 0;
    if (ok) {
        const V3 p = sp[S], d = sd[S];
        if constexpr (!WF) {
            // d sigma / d (px, py) = (p - centroid) / (N sigma)
            const double cx = moments[1] / cnt, cy = moments[2] / cnt;
            const double gx = (px - cx) / (cnt * obj), gy = (py - cy) / (cnt * obj);
            const double t = (A.L.detector_z - p.z) / d.z;
            pb = v3(gx, gy, 0);
            db = v3(t * gx, t * gy, 0);
            const double tb = gx * d.x + gy * d.y;
            pb.z += -tb / d.z;
            db.z += -tb * t / d.z;
        } else {
            // d wfe / d W = (W - mean_l) / (N wfe); W's last term
            // n_L (P - p) . d with P = (0, 0, detector_z)
            const double mu = moments[3 + 2 * wl] / moments[2 + 2 * wl];
            omega = (W - mu) / (cnt * obj);
            const double nL = nt[A.L.medium_after[S - 1]];
            pb = mul(d, -nL * omega);
            db = mul(v3(-p.x, -p.y, A.L.detector_z - p.z), nL * omega);
        }
    }
    for (int s = S - 1; s >= 0; --s) {
        double th[6] = {0, 0, 0, 0, 0, 0};
        double zth[ZK ? RPX_LENS_ZTERMS : 1];
        const double *C = ZK ? zrow(A.L, A.zc, s) : nullptr;
        const double Ru = A.L.zernike_radius[s];
#pragma unroll
        for (int k = 0; k < (ZK ? RPX_LENS_ZTERMS : 1); ++k) zth[k] = 0;
        if (ok) {
            const double *P = A.L.param[s];
            const V3 p = sp[s], d = sd[s];
            double t = 0;
            intersect(P, C, Ru, A.L.z[s], A.L.aperture[s] * A.L.aperture[s], p, d, t);   // recomputed
            const V3 q = add(p, mul(d, t));
            const double rho = q.x * q.x + q.y * q.y;
            const SurfEval e = surface_eval(P, rho, true);
            const ZEval z = zernike_eval(C, Ru, q.x, q.y, true);
            Refr r;
            refract(q, e.Sp, z, d, nt[A.L.medium_before[s]] / nt[A.L.medium_after[s]], r);
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
            // oriented unit normal -> unnormalised m = (-(2 x S' + Dx), -(2 y S' + Dy), 1)
            const V3 Nu = mul(r.N, r.sgn), Nub = mul(Nb, r.sgn);
            const V3 mb = mul(sub(Nub, mul(Nu, dot(Nub, Nu))), 1.0 / r.mlen);
            const double Spb = -2 * (q.x * mb.x + q.y * mb.y);
            const double Dxb = -mb.x, Dyb = -mb.y;
            qb.x += -2 * e.Sp * mb.x + Dxb * z.Dxx + Dyb * z.Dxy;
            qb.y += -2 * e.Sp * mb.y + Dxb * z.Dxy + Dyb * z.Dyy;
            const double rhob = Spb * e.Spp;
            for (int qn = 0; qn < 6; ++qn) th[qn] += Spb * e.dSp[qn];
            qb.x += 2 * q.x * rhob;
            qb.y += 2 * q.y * rhob;
            // intersection q = p + t d with G(t; p, d, param) = 0; the
            // segment's optical path n_before t adds to t's adjoint
            V3 pbar = qb;
            dbar = add(dbar, mul(qb, t));
            const double tbar = WF ? dot(qb, d) + omega * nt[A.L.medium_before[s]] : dot(qb, d);
            const V3 Gp = v3(2 * q.x * e.Sp + z.Dx, 2 * q.y * e.Sp + z.Dy, -1.0);
            const double Gt = dot(Gp, d);
            const double lam = -tbar / Gt;
            pbar = add(pbar, mul(Gp, lam));
            dbar = add(dbar, mul(Gp, lam * t));
            for (int qn = 0; qn < 6; ++qn) th[qn] += lam * e.dS[qn];
            if constexpr (ZK) {
              if (C) {
                // dG/dC_k = M_k(q); dm/dC_k = -(dM_k/dx, dM_k/dy, 0)
                const double inv = 1.0 / Ru;
                double xp[ZDEG + 1], yp[ZDEG + 1];
                zpowers(q.x * inv, xp);
                zpowers(q.y * inv, yp);
                int k = 0;
#pragma unroll
                for (int dg = 0; dg <= ZDEG; ++dg) {
#pragma unroll
                    for (int a = dg; a >= 0; --a, ++k) {
                        const int b = dg - a;
                        double v = lam * xp[a] * yp[b];
                        if (a >= 1) v += Dxb * a * xp[a - 1] * yp[b] * inv;
                        if (b >= 1) v += Dyb * b * xp[a] * yp[b - 1] * inv;
                        zth[k] = v;
                    }
                }
              }
            }
            pb = pbar;
            db = dbar;
        }
        for (int qn = 0; qn < 6; ++qn) {
            const double v = wave_sum(th[qn]);
            if (lane == 0) gsum[wave][s * 6 + qn] = v;
        }
        if constexpr (ZK) {
            if (C) {
                for (int k = 0; k < RPX_LENS_ZTERMS; ++k) {
                    const double v = wave_sum(zth[k]);
                    if (lane == 0) gsum[wave][zbase + s * RPX_LENS_ZTERMS + k] = v;
                }
            } else if (lane == 0) {
                for (int k = 0; k < RPX_LENS_ZTERMS; ++k) gsum[wave][zbase + s * RPX_LENS_ZTERMS + k] = 0.0;
            }
        } else if (lane == 0) {
            for (int k = 0; k < RPX_LENS_ZTERMS; ++k) gsum[wave][zbase + s * RPX_LENS_ZTERMS + k] = 0.0;
        }
    }
    __syncthreads();
    const int width = S * (6 + RPX_LENS_ZTERMS);
    for (int k = threadIdx.x; k < width; k += LB) {
        double v = 0;
        for (int w = 0; w < LB / 64; ++w) v += gsum[w][k];
        partials[(uint64_t)blockIdx.x * width + k] = v;
    }
}

// Without distortions the adjoint fits 5 waves per SIMD (96 VGPRs).
template <bool WF>
__global__ __launch_bounds__(LB) __attribute__((amdgpu_waves_per_eu(5))) void k_lens_adjoint(
    LensArgs A, const double *moments, double *partials) {
    lens_adjoint<false, WF>(A, moments, partials);
}

template <bool WF>
__global__ __launch_bounds__(LB) void k_lens_adjoint_zernike(LensArgs A, const double *moments, double *partials) {
    lens_adjoint<true, WF>(A, moments, partials);
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

rpx_status_t check_objective(const rpx_lens_t *L, const double *zcart, uint64_t n_wl, int32_t objective) {
    if (objective != RPX_LENS_SPOT_RMS && objective != RPX_LENS_WAVEFRONT_RMS)
        return rpx_set_error(RPX_E_INVALID, "unknown lens objective");
    if (objective == RPX_LENS_WAVEFRONT_RMS && n_wl > RPX_LENS_MAX_WFE_WAVELENGTHS)
        return rpx_set_error(RPX_E_UNSUPPORTED, "the wavefront objective takes at most 16 wavelengths");
    for (int s = 0; s < L->n_surfaces; ++s)
        if (L->zernike_radius[s] > 0 && !zcart)
            return rpx_set_error(RPX_E_INVALID, "a surface has a distortion but the monomial table is null");
    return RPX_OK;
}

bool has_distortion(const rpx_lens_t *L, const double *zcart) {
    if (!zcart) return false;
    for (int s = 0; s < L->n_surfaces; ++s)
        if (L->zernike_radius[s] > 0) return true;
    return false;
}

uint64_t lens_width(int32_t n_surfaces) {
    return (uint64_t)std::max(RPX_LENS_MOMENTS, (6 + RPX_LENS_ZTERMS) * std::max(1, (int)n_surfaces));
}

rpx_status_t lens_eval(const rpx_lens_t *lens, const double *zcart, const rpx_rays_t *rays, uint64_t n_rays,
                       const double *n_table, uint64_t n_wavelengths, int32_t objective, double *moments,
                       void *workspace, uint64_t workspace_bytes, hipStream_t s) {
    if (!moments) return rpx_set_error(RPX_E_INVALID, "moments is null");
    if (!workspace || workspace_bytes < rpx_lens_workspace_bytes(n_rays, lens->n_surfaces))
        return rpx_set_error(RPX_E_INVALID, "workspace too small");
    const uint64_t nb = std::max<uint64_t>((n_rays + LB - 1) / LB, 1);
    LensArgs A{*lens, *rays, n_rays, n_table, n_wavelengths, zcart, objective};
    double *part = (double *)workspace;
    const bool zk = has_distortion(lens, zcart), wf = objective == RPX_LENS_WAVEFRONT_RMS;
    auto forward = zk ? (wf ? k_lens_forward<true, true> : k_lens_forward<true, false>)
                      : (wf ? k_lens_forward<false, true> : k_lens_forward<false, false>);
    if (objective == RPX_LENS_SPOT_RMS) {
        hipLaunchKernelGGL(forward, dim3((unsigned)nb), dim3(LB), 0, s, A, 0, (const double *)moments, part);
        RPX_HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_lens_reduce, dim3(4), dim3(256), 0, s, (const double *)part, nb, 4, moments);
        RPX_HIP_TRY(hipGetLastError());
        return RPX_OK;
    }
    const int w0 = 2 * (int)n_wavelengths;
    hipLaunchKernelGGL(forward, dim3((unsigned)nb), dim3(LB), 0, s, A, 0, (const double *)moments, part);
    RPX_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_lens_reduce, dim3(w0), dim3(256), 0, s, (const double *)part, nb, w0, moments + 2);
    RPX_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(forward, dim3((unsigned)nb), dim3(LB), 0, s, A, 1, (const double *)moments, part);
    RPX_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_lens_reduce, dim3(2), dim3(256), 0, s, (const double *)part, nb, 2, moments);
    RPX_HIP_TRY(hipGetLastError());
    return RPX_OK;
}

rpx_status_t lens_grad(const rpx_lens_t *lens, const double *zcart, const rpx_rays_t *rays, uint64_t n_rays,
                       const double *n_table, uint64_t n_wavelengths, int32_t objective, const double *moments,
                       double *grad, int columns, void *workspace, uint64_t workspace_bytes, hipStream_t s) {
    if (!moments || !grad) return rpx_set_error(RPX_E_INVALID, "moments or grad is null");
    if (!workspace || workspace_bytes < rpx_lens_workspace_bytes(n_rays, lens->n_surfaces))
        return rpx_set_error(RPX_E_INVALID, "workspace too small");
    const uint64_t nb = std::max<uint64_t>((n_rays + LB - 1) / LB, 1);
    LensArgs A{*lens, *rays, n_rays, n_table, n_wavelengths, zcart, objective};
    const int width = (6 + RPX_LENS_ZTERMS) * lens->n_surfaces;
    const bool zk = has_distortion(lens, zcart), wf = objective == RPX_LENS_WAVEFRONT_RMS;
    auto adjoint = zk ? (wf ? k_lens_adjoint_zernike<true> : k_lens_adjoint_zernike<false>)
                      : (wf ? k_lens_adjoint<true> : k_lens_adjoint<false>);
    hipLaunchKernelGGL(adjoint, dim3((unsigned)nb), dim3(LB), 0, s, A, moments, (double *)workspace);
    RPX_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_lens_reduce, dim3(columns), dim3(256), 0, s, (const double *)workspace, nb, width, grad);
    RPX_HIP_TRY(hipGetLastError());
    return RPX_OK;
}

}  // namespace

extern "C" uint64_t rpx_lens_workspace_bytes(uint64_t n_rays, int32_t n_surfaces) {
    const uint64_t nb = (n_rays + LB - 1) / LB;
    return std::max<uint64_t>(nb, 1) * lens_width(n_surfaces) * 8;
}

extern "C" rpx_status_t rpx_lens_eval(const rpx_lens_t *lens, const double *zcart, const rpx_rays_t *rays,
                                      uint64_t n_rays, const double *n_table, uint64_t n_wavelengths,
                                      int32_t objective, double *moments, void *workspace, uint64_t workspace_bytes,
                                      void *stream) {
    rpx_status_t st = check_lens(lens, rays, n_rays, n_table, n_wavelengths);
    if (!st) st = check_objective(lens, zcart, n_wavelengths, objective);
    if (st) return st;
    return lens_eval(lens, zcart, rays, n_rays, n_table, n_wavelengths, objective, moments, workspace,
                     workspace_bytes, (hipStream_t)stream);
}

extern "C" rpx_status_t rpx_lens_grad(const rpx_lens_t *lens, const double *zcart, const rpx_rays_t *rays,
                                      uint64_t n_rays, const double *n_table, uint64_t n_wavelengths,
                                      int32_t objective, const double *moments, double *grad, void *workspace,
                                      uint64_t workspace_bytes, void *stream) {
    rpx_status_t st = check_lens(lens, rays, n_rays, n_table, n_wavelengths);
    if (!st) st = check_objective(lens, zcart, n_wavelengths, objective);
    if (st) return st;
    return lens_grad(lens, zcart, rays, n_rays, n_table, n_wavelengths, objective, moments, grad,
                     (6 + RPX_LENS_ZTERMS) * lens->n_surfaces, workspace, workspace_bytes, (hipStream_t)stream);
}

extern "C" rpx_status_t rpx_lens_spot(const rpx_lens_t *lens, const rpx_rays_t *rays, uint64_t n_rays,
                                      const double *n_table, uint64_t n_wavelengths, double *moments, void *workspace,
                                      uint64_t workspace_bytes, void *stream) {
    rpx_status_t st = check_lens(lens, rays, n_rays, n_table, n_wavelengths);
    if (st) return st;
    rpx_lens_t L = *lens;
    for (int s = 0; s < RPX_LENS_MAX_SURFACES; ++s) L.zernike_radius[s] = 0;
    return lens_eval(&L, nullptr, rays, n_rays, n_table, n_wavelengths, RPX_LENS_SPOT_RMS, moments, workspace,
                     workspace_bytes, (hipStream_t)stream);
}

extern "C" rpx_status_t rpx_lens_spot_grad(const rpx_lens_t *lens, const rpx_rays_t *rays, uint64_t n_rays,
                                           const double *n_table, uint64_t n_wavelengths, const double *moments,
                                           double *grad, void *workspace, uint64_t workspace_bytes, void *stream) {
    rpx_status_t st = check_lens(lens, rays, n_rays, n_table, n_wavelengths);
    if (st) return st;
    rpx_lens_t L = *lens;
    for (int s = 0; s < RPX_LENS_MAX_SURFACES; ++s) L.zernike_radius[s] = 0;
    return lens_grad(&L, nullptr, rays, n_rays, n_table, n_wavelengths, RPX_LENS_SPOT_RMS, moments, grad,
                     6 * lens->n_surfaces, workspace, workspace_bytes, (hipStream_t)stream);
}
