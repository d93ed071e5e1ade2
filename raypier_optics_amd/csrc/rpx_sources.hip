// Ray and gausslet sources generated in HBM.
//
//  * rpx_lattice_select / rpx_lattice_emit -- the hexagonal lattices of
//    HexagonalRayFieldSource (sources.py:1035-1107), CollimatedGaussletSource
//    (gausslet_sources.py:386-457) and GaussianPointSource
//    (gausslet_sources.py:516-566 with make_hexagonal_grid,
//    core/gausslets.py:55-80).  The lattice indices run over
//    [-nsteps-1, nsteps+1]^2 in numpy.meshgrid order (label = row * side +
//    column, the row giving j); the selected points keep that order.
//  * rpx_config_parabasal_rays -- GaussletCollection.config_parabasal_rays
//    (ctracer.pyx:1430-1481).
//  * rpx_project_to_plane -- GaussletCollection.project_to_plane
//    (ctracer.pyx:1391-1420).
//  * rpx_lagrange_invariant -- GaussletCollection.lagrange_invariant
//    (ctracer.pyx:1347-1389).
//
// Every expression follows the reference's evaluation order (numpy
// elementwise, or the Cython vector helpers) and is compiled without
// contraction; the exponential of the amplitude weight is the device's.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <string>
#include "rpx_internal.h"
#include "rpx_host.h"
#include "rpx_gen.h"

using namespace rpx;

namespace {

struct LatticeWs {
    int32_t *backmap;      // label -> output index, or -1
    uint64_t *tile_status;
    double *partials;      // per tile: sum of power * g^2 over its selected points
    double *power;         // total of partials
    GenControl *ctl;
};

uint64_t n_labels(int32_t nsteps) {
    const uint64_t side = 2 * (uint64_t)nsteps + 3;
    return side * side;
}

uint64_t lattice_bytes(int32_t nsteps) {
    const uint64_t nl = n_labels(nsteps), nb = (nl + RPX_BLOCK - 1) / RPX_BLOCK;
    return align256(4 * nl) + 2 * align256(8 * nb) + align256(8) + align256(sizeof(GenControl));
}

LatticeWs carve_lattice(void *ws, int32_t nsteps) {
    const uint64_t nl = n_labels(nsteps), nb = (nl + RPX_BLOCK - 1) / RPX_BLOCK;
    char *p = (char *)ws;
    LatticeWs w;
    w.backmap = (int32_t *)p;
    p += align256(4 * nl);
    w.tile_status = (uint64_t *)p;
    p += align256(8 * nb);
    w.partials = (double *)p;
    p += align256(8 * nb);
    w.power = (double *)p;
    p += align256(8);
    w.ctl = (GenControl *)p;
    return w;
}

struct LatticePoint {
    double vi, vj;   // lattice indices
    double r;        // ri (ray field) or kx^2 + ky^2 (point source)
    double kx, ky;   // point source only
    bool sel;
};

__device__ __forceinline__ LatticePoint lattice_point(const rpx_lattice_t &L, uint64_t label) {
    const uint64_t side = 2 * (uint64_t)L.nsteps + 3;
    const int64_t row = (int64_t)(label / side), col = (int64_t)(label % side);
    LatticePoint q;
    const int64_t ii = col - L.nsteps - 1, jj = row - L.nsteps - 1;
    q.vi = (double)ii;
    q.vj = (double)jj;
    if (L.kind == RPX_LATTICE_RAY_FIELD) {
        // ri = (vi + sinV*vj)**2 + (cosV*vj**2), vj**2 an integer square
        const double a = q.vi + L.sin30 * q.vj;
        const double b = L.cos30 * (double)(jj * jj);
        q.r = a * a + b;
        q.sel = q.r < L.limit;
        q.kx = q.ky = 0.0;
    } else {
        // x = (vi + sinV*vj)*spacing, y = cosV*vj*spacing, r2 = x**2 + y**2
        q.kx = (q.vi + L.sin30 * q.vj) * L.spacing;
        q.ky = (L.cos30 * q.vj) * L.spacing;
        q.r = q.kx * q.kx + q.ky * q.ky;
        q.sel = q.r < L.limit;
    }
    return q;
}

__device__ __forceinline__ double lattice_weight(const rpx_lattice_t &L, double r) {
    return exp(-((r * L.weight_num) / L.weight_den));
}

}  // namespace

// Numbers the selected lattice points in label order (decoupled look-back
// over tiles of labels) and sums power * g^2 per tile.
__global__ __launch_bounds__(RPX_BLOCK) void k_lattice_select(rpx_lattice_t L, uint64_t n_lab, LatticeWs w) {
    __shared__ uint32_t lds_wave[RPX_BLOCK / 64];
    __shared__ double lds_sum[RPX_BLOCK];
    __shared__ uint32_t lds_tile;
    __shared__ uint64_t lds_excl;
    if (threadIdx.x == 0) lds_tile = atomicAdd(&w.ctl->tile_counter, 1u);
    __syncthreads();
    const uint32_t tile = lds_tile;
    const uint64_t label = (uint64_t)tile * RPX_BLOCK + threadIdx.x;
    bool sel = false;
    double q = 0.0;
    if (label < n_lab) {
        const LatticePoint p = lattice_point(L, label);
        sel = p.sel;
        if (sel && L.power != 0.0) {
            const double g = lattice_weight(L, p.r);
            q = L.power * (g * g);
        }
    }
    uint32_t block_total;
    const uint32_t local = block_scan_small(sel ? 1u : 0u, lds_wave, &block_total);
    if (threadIdx.x < 64) {
        const uint64_t excl = lookback(w.tile_status, tile, 0, block_total, &w.ctl->error);
        if (threadIdx.x == 0) lds_excl = excl;
    }
    lds_sum[threadIdx.x] = q;
    __syncthreads();
    for (int s = RPX_BLOCK / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) lds_sum[threadIdx.x] = lds_sum[threadIdx.x] + lds_sum[threadIdx.x + s];
        __syncthreads();
    }
    if (label < n_lab) w.backmap[label] = sel ? (int32_t)(lds_excl + local) : -1;
    if (threadIdx.x == 0) {
        w.partials[tile] = lds_sum[0];
        if ((uint64_t)tile == (n_lab - 1) / RPX_BLOCK) w.ctl->children = lds_excl + block_total;
    }
}

// Sum of the tile partials in a fixed order (one workgroup).
__global__ __launch_bounds__(256) void k_sum_partials(const double *partials, uint64_t n, double *out) {
    __shared__ double s[256];
    double acc = 0.0;
    for (uint64_t k = threadIdx.x; k < n; k += 256) acc = acc + partials[k];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

// Writes the ray (base ray and parabasal copies for gausslets, as
// GaussletCollection.from_rays, ctracer.pyx:1321-1345) of every selected
// label, and its six neighbours for the ray field (sources.py:1063-1085).
__global__ __launch_bounds__(256) void k_lattice_emit(rpx_lattice_t L, uint64_t n_lab, const int32_t *backmap,
                                                      rpx_rays_t out, int32_t *neighbours) {
    const uint64_t label = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (label >= n_lab) return;
    const int32_t pos32 = backmap[label];
    if (pos32 < 0) return;
    const uint64_t pos = (uint64_t)pos32;
    if (pos >= out.capacity) return;
    const LatticePoint p = lattice_point(L, label);
    const double g = lattice_weight(L, p.r);
    double o[3], d[3], e[3];
    if (L.kind == RPX_LATTICE_RAY_FIELD) {
        // offsets = xi*d1 + yj*d2; offsets *= spacing; offsets += origin
        for (int k = 0; k < 3; ++k) {
            const double off = p.vi * L.d1[k] + p.vj * L.d2[k];
            o[k] = off * L.spacing + L.origin[k];
            d[k] = L.direction[k];
            e[k] = L.E_vector[k];
        }
    } else {
        // directions = (kx*d1 + ky*d2 + kz*direction)/k_abs, normalised;
        // E_vector = cross(direction, d2)
        const double kz = sqrt(L.k_abs_sq - p.r);
        for (int k = 0; k < 3; ++k) {
            const double a = p.kx * L.d1[k] + p.ky * L.d2[k];
            d[k] = (a + kz * L.direction[k]) / L.k_abs;
            o[k] = L.origin[k];
        }
        const double mag = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
        for (int k = 0; k < 3; ++k) d[k] = d[k] / mag;
        e[0] = d[1] * L.d2[2] - d[2] * L.d2[1];
        e[1] = d[2] * L.d2[0] - d[0] * L.d2[2];
        e[2] = d[0] * L.d2[1] - d[1] * L.d2[0];
    }
    double *f = out.f64;
    const uint64_t c = out.capacity;
    f[RPX_OX * c + pos] = o[0];  f[RPX_OY * c + pos] = o[1];  f[RPX_OZ * c + pos] = o[2];
    f[RPX_DX * c + pos] = d[0];  f[RPX_DY * c + pos] = d[1];  f[RPX_DZ * c + pos] = d[2];
    f[RPX_NX * c + pos] = 0.0;   f[RPX_NY * c + pos] = 1.0;   f[RPX_NZ * c + pos] = 0.0;
    f[RPX_EX * c + pos] = e[0];  f[RPX_EY * c + pos] = e[1];  f[RPX_EZ * c + pos] = e[2];
    f[RPX_N_RE * c + pos] = 1.0;
    f[RPX_N_IM * c + pos] = 0.0;
    // complex amplitude times the real weight, as numpy's complex product
    f[RPX_E1_RE * c + pos] = L.E1[0] * g - L.E1[1] * 0.0;
    f[RPX_E1_IM * c + pos] = L.E1[1] * g + L.E1[0] * 0.0;
    f[RPX_E2_RE * c + pos] = L.E2[0] * g - L.E2[1] * 0.0;
    f[RPX_E2_IM * c + pos] = L.E2[1] * g + L.E2[0] * 0.0;
    f[RPX_LENGTH * c + pos] = 0.0;
    f[RPX_PHASE * c + pos] = 0.0;
    f[RPX_ACCUM * c + pos] = 0.0;
    uint32_t *u = out.u32;
    u[RPX_WL_IDX * c + pos] = 0;
    u[RPX_PARENT_IDX * c + pos] = 0;
    u[RPX_END_FACE_IDX * c + pos] = 0;
    u[RPX_RAY_IDENT * c + pos] = L.ray_ident_index ? (uint32_t)pos : 0u;
    u[RPX_RAY_TYPE_ID * c + pos] = L.ray_type_id;
    if (out.para) {
        for (int j = 0; j < RPX_N_PARA; ++j) {
            const uint64_t b = (uint64_t)j * RPX_PARA_FIELDS;
            for (int k = 0; k < 3; ++k) {
                out.para[(b + k) * c + pos] = o[k];
                out.para[(b + 3 + k) * c + pos] = d[k];
            }
            out.para[(b + 6) * c + pos] = 0.0;
            out.para[(b + 7) * c + pos] = 1.0;
            out.para[(b + 8) * c + pos] = 0.0;
            out.para[(b + 9) * c + pos] = 0.0;
        }
    }
    if (neighbours) {
        // neighbour k of label (r, c): 0 (r+1, c), 1 (r, c+1), 2 (r-1, c+1),
        // 3 (r-1, c), 4 (r, c-1), 5 (r+1, c-1)
        const int64_t side = 2 * (int64_t)L.nsteps + 3;
        const int64_t r = (int64_t)(label / side), cc = (int64_t)(label % side);
        const int dr[6] = {1, 0, -1, -1, 0, 1}, dc[6] = {0, 1, 1, 0, -1, -1};
        for (int k = 0; k < 6; ++k) {
            const int64_t r2 = r + dr[k], c2 = cc + dc[k];
            int32_t v = -1;
            if (r2 >= 0 && r2 < side && c2 >= 0 && c2 < side) v = backmap[r2 * side + c2];
            neighbours[pos * 6 + k] = v;
        }
    }
}

extern "C" uint64_t rpx_lattice_workspace_bytes(int32_t nsteps) {
    if (nsteps < 0) return 0;
    return lattice_bytes(nsteps);
}

static rpx_status_t check_lattice(const rpx_lattice_t *L) {
    if (!L) return rpx_set_error(RPX_E_INVALID, "lattice is null");
    if (L->kind != RPX_LATTICE_RAY_FIELD && L->kind != RPX_LATTICE_POINT_SOURCE)
        return rpx_set_error(RPX_E_INVALID, "unknown lattice kind");
    if (L->nsteps < 0 || L->nsteps > 16384) return rpx_set_error(RPX_E_UNSUPPORTED, "nsteps outside 0..16384");
    return RPX_OK;
}

extern "C" rpx_status_t rpx_lattice_select(const rpx_lattice_t *L, uint64_t *n_points, double *weight_power,
                                           void *workspace, uint64_t workspace_bytes, void *stream_) {
    if (rpx_status_t st = check_lattice(L)) return st;
    if (!n_points) return rpx_set_error(RPX_E_INVALID, "n_points is null");
    *n_points = 0;
    if (!workspace || workspace_bytes < lattice_bytes(L->nsteps))
        return rpx_set_error(RPX_E_INVALID, "workspace too small");
    hipStream_t stream = (hipStream_t)stream_;
    const uint64_t nl = n_labels(L->nsteps), nb = (nl + RPX_BLOCK - 1) / RPX_BLOCK;
    LatticeWs w = carve_lattice(workspace, L->nsteps);
    hipLaunchKernelGGL(k_clear, dim3((unsigned)((std::max<uint64_t>(nb, RPX_SHARDS * RPX_LDS_FACES) + 255) / 256)),
                       dim3(256), 0, stream, w.tile_status, nb, w.ctl);
    RPX_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_lattice_select, dim3((unsigned)nb), dim3(RPX_BLOCK), 0, stream, *L, nl, w);
    RPX_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, stream, (const double *)w.partials, nb, w.power);
    RPX_HIP_TRY(hipGetLastError());
    GenControl ctl;
    double power = 0.0;
    RPX_HIP_TRY(hipMemcpyAsync(&ctl, w.ctl, RPX_CTL_HEAD, hipMemcpyDeviceToHost, stream));
    RPX_HIP_TRY(hipMemcpyAsync(&power, w.power, sizeof(double), hipMemcpyDeviceToHost, stream));
    RPX_HIP_TRY(hipStreamSynchronize(stream));
    if (ctl.error) return rpx_set_error(RPX_E_HIP, "lattice scan: a look-back wait exceeded its poll limit");
    *n_points = ctl.children;
    if (weight_power) *weight_power = power;
    return RPX_OK;
}

extern "C" rpx_status_t rpx_lattice_emit(const rpx_lattice_t *L, rpx_rays_t *out, int32_t *neighbours,
                                         void *workspace, uint64_t workspace_bytes, void *stream_) {
    if (rpx_status_t st = check_lattice(L)) return st;
    if (!out) return rpx_set_error(RPX_E_INVALID, "output rays are null");
    if (out->capacity > 0 && (!out->f64 || !out->u32)) return rpx_set_error(RPX_E_INVALID, "output has null fields");
    if (!workspace || workspace_bytes < lattice_bytes(L->nsteps))
        return rpx_set_error(RPX_E_INVALID, "workspace too small");
    hipStream_t stream = (hipStream_t)stream_;
    const uint64_t nl = n_labels(L->nsteps);
    LatticeWs w = carve_lattice(workspace, L->nsteps);
    GenControl ctl;
    RPX_HIP_TRY(read_control(ctl, w.ctl, stream));
    if (ctl.children > out->capacity)
        return rpx_set_error(RPX_E_CAPACITY, "output capacity " + std::to_string(out->capacity) + " below " +
                                                 std::to_string(ctl.children) + " lattice points");
    hipLaunchKernelGGL(k_lattice_emit, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, stream, *L, nl,
                       (const int32_t *)w.backmap, *out, neighbours);
    RPX_HIP_TRY(hipGetLastError());
    RPX_HIP_TRY(hipStreamSynchronize(stream));
    return RPX_OK;
}

// ------------------------------------------------------ parabasal rays
__device__ __forceinline__ V3 norm3(V3 a) {
    const double m = sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return v3(a.x / m, a.y / m, a.z / m);
}
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

__global__ __launch_bounds__(256) void k_config_parabasal(rpx_rays_t g, uint64_t n, const double *wl, uint64_t n_wl,
                                                          double radius, double working_dist, double c0, double s0,
                                                          double c1, double s1, double c2, double s2,
                                                          unsigned long long *err) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *f = g.f64;
    const uint64_t c = g.capacity;
    const V3 base_d = norm3(v3(f[RPX_DX * c + i], f[RPX_DY * c + i], f[RPX_DZ * c + i]));
    const V3 origin = v3(f[RPX_OX * c + i], f[RPX_OY * c + i], f[RPX_OZ * c + i]);
    const V3 normal = v3(f[RPX_NX * c + i], f[RPX_NY * c + i], f[RPX_NZ * c + i]);
    const double length = f[RPX_LENGTH * c + i];
    const uint32_t wi = g.u32[RPX_WL_IDX * c + i];
    if (wi >= n_wl) {
        atomicAdd(err, 1ull);
        return;
    }
    const V3 o0 = base_d.x > base_d.y ? v3(0.0, 1.0, 0.0) : v3(1.0, 0.0, 0.0);
    const V3 d1 = norm3(cross3(base_d, o0));
    const V3 d2 = norm3(cross3(base_d, d1));
    const double theta0 = wl[wi] / (M_PI * radius * 1000.0);
    const double cs[3] = {c0, c1, c2}, sn[3] = {s0, s1, s2};
    for (int q = 0; q < 3; ++q) {
        const int j = 2 * q;
        // o = d1*(radius*cos) + d2*(radius*sin) + origin + base_d*working_dist
        V3 o = add(mul(d1, radius * cs[q]), mul(d2, radius * sn[q]));
        o = add(o, origin);
        o = add(o, mul(base_d, working_dist));
        const V3 d = add(mul(d1, -theta0 * sn[q]), mul(d2, theta0 * cs[q]));
        const V3 da = add(base_d, d), db = sub(base_d, d);
        const V3 na = norm3(da), nb = norm3(db);
        const V3 oa = sub(o, mul(da, working_dist)), ob = sub(o, mul(db, working_dist));
        const V3 dirs[2] = {na, nb}, orgs[2] = {oa, ob};
        for (int h = 0; h < 2; ++h) {
            const uint64_t b = (uint64_t)(j + h) * RPX_PARA_FIELDS;
            g.para[(b + 0) * c + i] = orgs[h].x;
            g.para[(b + 1) * c + i] = orgs[h].y;
            g.para[(b + 2) * c + i] = orgs[h].z;
            g.para[(b + 3) * c + i] = dirs[h].x;
            g.para[(b + 4) * c + i] = dirs[h].y;
            g.para[(b + 5) * c + i] = dirs[h].z;
            g.para[(b + 6) * c + i] = normal.x;
            g.para[(b + 7) * c + i] = normal.y;
            g.para[(b + 8) * c + i] = normal.z;
            g.para[(b + 9) * c + i] = length;
        }
    }
}

extern "C" rpx_status_t rpx_config_parabasal_rays(rpx_rays_t *gausslets, uint64_t n, const double *wavelengths,
                                                  uint64_t n_wavelengths, double radius, double working_dist,
                                                  const double *angle_cos_sin, void *workspace,
                                                  uint64_t workspace_bytes, void *stream_) {
    if (!gausslets) return rpx_set_error(RPX_E_INVALID, "gausslets are null");
    if (gausslets->capacity < n) return rpx_set_error(RPX_E_INVALID, "gausslets: capacity below count");
    if (n == 0) return RPX_OK;
    if (!gausslets->f64 || !gausslets->u32 || !gausslets->para)
        return rpx_set_error(RPX_E_INVALID, "gausslets have null fields");
    if (!wavelengths && n_wavelengths) return rpx_set_error(RPX_E_INVALID, "wavelengths are null");
    if (!angle_cos_sin) return rpx_set_error(RPX_E_INVALID, "angle table is null");
    if (!workspace || workspace_bytes < 8) return rpx_set_error(RPX_E_INVALID, "workspace too small");
    hipStream_t stream = (hipStream_t)stream_;
    unsigned long long *err = (unsigned long long *)workspace;
    RPX_HIP_TRY(hipMemsetAsync(err, 0, sizeof(*err), stream));
    const double *a = angle_cos_sin;
    hipLaunchKernelGGL(k_config_parabasal, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, *gausslets, n,
                       wavelengths, n_wavelengths, radius, working_dist, a[0], a[1], a[2], a[3], a[4], a[5], err);
    RPX_HIP_TRY(hipGetLastError());
    unsigned long long bad = 0;
    RPX_HIP_TRY(hipMemcpyAsync(&bad, err, sizeof(bad), hipMemcpyDeviceToHost, stream));
    RPX_HIP_TRY(hipStreamSynchronize(stream));
    if (bad) return rpx_set_error(RPX_E_INVALID, std::to_string(bad) + " gausslets have a wavelength_idx outside "
                                                 "the wavelength list");
    return RPX_OK;
}

// ------------------------------------------------------ plane projection
__global__ __launch_bounds__(256) void k_project_to_plane(rpx_rays_t g, uint64_t n, V3 o, V3 d) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double *f = g.f64;
    const uint64_t c = g.capacity;
    const V3 ro = v3(f[RPX_OX * c + i], f[RPX_OY * c + i], f[RPX_OZ * c + i]);
    const V3 rd = v3(f[RPX_DX * c + i], f[RPX_DY * c + i], f[RPX_DZ * c + i]);
    const double a = dot(sub(o, ro), d) / dot(rd, d);
    const V3 no = add(ro, mul(rd, a));
    f[RPX_OX * c + i] = no.x;
    f[RPX_OY * c + i] = no.y;
    f[RPX_OZ * c + i] = no.z;
    f[RPX_ACCUM * c + i] = f[RPX_ACCUM * c + i] + f[RPX_N_RE * c + i] * a;
    if (g.para) {
        for (int j = 0; j < RPX_N_PARA; ++j) {
            const uint64_t b = (uint64_t)j * RPX_PARA_FIELDS;
            const V3 po = v3(g.para[(b + 0) * c + i], g.para[(b + 1) * c + i], g.para[(b + 2) * c + i]);
            const V3 pd = v3(g.para[(b + 3) * c + i], g.para[(b + 4) * c + i], g.para[(b + 5) * c + i]);
            const double ap = dot(sub(o, po), d) / dot(pd, d);
            const V3 np = add(po, mul(pd, ap));
            g.para[(b + 0) * c + i] = np.x;
            g.para[(b + 1) * c + i] = np.y;
            g.para[(b + 2) * c + i] = np.z;
        }
    }
}

extern "C" rpx_status_t rpx_project_to_plane(rpx_rays_t *rays, uint64_t n, const double *origin,
                                             const double *direction, void *stream_) {
    if (!rays) return rpx_set_error(RPX_E_INVALID, "rays are null");
    if (rays->capacity < n) return rpx_set_error(RPX_E_INVALID, "rays: capacity below count");
    if (n == 0) return RPX_OK;
    if (!rays->f64) return rpx_set_error(RPX_E_INVALID, "rays have null fields");
    if (!origin || !direction) return rpx_set_error(RPX_E_INVALID, "plane is null");
    hipStream_t stream = (hipStream_t)stream_;
    // d = norm_(direction) (ctracer.pyx:1407), as the reference, on the host
    const double m = std::sqrt(direction[0] * direction[0] + direction[1] * direction[1] + direction[2] * direction[2]);
    V3 d, o;
    d.x = direction[0] / m;
    d.y = direction[1] / m;
    d.z = direction[2] / m;
    o.x = origin[0];
    o.y = origin[1];
    o.z = origin[2];
    hipLaunchKernelGGL(k_project_to_plane, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, *rays, n, o, d);
    RPX_HIP_TRY(hipGetLastError());
    RPX_HIP_TRY(hipStreamSynchronize(stream));
    return RPX_OK;
}

// ------------------------------------------------------ Lagrange invariant
// One gausslet per lane.  Each parabasal ray's offset from the base origin
// and its direction are taken in the frame (a1, a2) = (E / |E|, a1 x d);
// the six pairs (0,5) (1,2) (3,4) (1,4) (0,3) (2,5) add (h_i.u_j - h_j.u_i)^2
// in that order, and the result is 1000 sqrt(v / 6) / wavelength.
__global__ __launch_bounds__(256) void k_lagrange_invariant(rpx_rays_t g, uint64_t n, const double *wl,
                                                            uint64_t n_wl, double *out, unsigned long long *err) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *f = g.f64;
    const uint64_t c = g.capacity;
    const uint32_t wi = g.u32[RPX_WL_IDX * c + i];
    if (wi >= n_wl) {
        atomicAdd(err, 1ull);
        out[i] = 0.0;
        return;
    }
    const V3 a1 = norm3(v3(f[RPX_EX * c + i], f[RPX_EY * c + i], f[RPX_EZ * c + i]));
    const V3 a2 = cross3(a1, v3(f[RPX_DX * c + i], f[RPX_DY * c + i], f[RPX_DZ * c + i]));
    const V3 ro = v3(f[RPX_OX * c + i], f[RPX_OY * c + i], f[RPX_OZ * c + i]);
    double hx[RPX_N_PARA], hy[RPX_N_PARA], ux[RPX_N_PARA], uy[RPX_N_PARA];
#pragma unroll
    for (int j = 0; j < RPX_N_PARA; ++j) {
        const uint64_t b = (uint64_t)j * RPX_PARA_FIELDS;
        const V3 o = sub(v3(g.para[(b + 0) * c + i], g.para[(b + 1) * c + i], g.para[(b + 2) * c + i]), ro);
        const V3 d = v3(g.para[(b + 3) * c + i], g.para[(b + 4) * c + i], g.para[(b + 5) * c + i]);
        hx[j] = dot(o, a1);
        hy[j] = dot(o, a2);
        ux[j] = dot(d, a1);
        uy[j] = dot(d, a2);
    }
    // dot products of the planar (z = 0) vectors, summed as x + y + 0
    auto term = [&](int a, int b) {
        const double t = ((hx[a] * ux[b] + hy[a] * uy[b]) + 0.0) - ((hx[b] * ux[a] + hy[b] * uy[a]) + 0.0);
        return t * t;
    };
    double v = 0.0;
    v += term(0, 5);
    v += term(1, 2);
    v += term(3, 4);
    v += term(1, 4);
    v += term(0, 3);
    v += term(2, 5);
    out[i] = 1000.0 * sqrt(v / 6.0) / wl[wi];
}

extern "C" rpx_status_t rpx_lagrange_invariant(const rpx_rays_t *gausslets, uint64_t n, const double *wavelengths,
                                               uint64_t n_wavelengths, double *out, void *workspace,
                                               uint64_t workspace_bytes, void *stream_) {
    if (!gausslets) return rpx_set_error(RPX_E_INVALID, "gausslets are null");
    if (gausslets->capacity < n) return rpx_set_error(RPX_E_INVALID, "gausslets: capacity below count");
    if (n == 0) return RPX_OK;
    if (!gausslets->f64 || !gausslets->u32 || !gausslets->para)
        return rpx_set_error(RPX_E_INVALID, "gausslets have null fields");
    if (!out) return rpx_set_error(RPX_E_INVALID, "output is null");
    if (!wavelengths && n_wavelengths) return rpx_set_error(RPX_E_INVALID, "wavelengths are null");
    if (!workspace || workspace_bytes < 8) return rpx_set_error(RPX_E_INVALID, "workspace too small");
    hipStream_t stream = (hipStream_t)stream_;
    unsigned long long *err = (unsigned long long *)workspace;
    RPX_HIP_TRY(hipMemsetAsync(err, 0, sizeof(*err), stream));
    hipLaunchKernelGGL(k_lagrange_invariant, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, *gausslets, n,
                       wavelengths, n_wavelengths, out, err);
    RPX_HIP_TRY(hipGetLastError());
    unsigned long long bad = 0;
    RPX_HIP_TRY(hipMemcpyAsync(&bad, err, sizeof(bad), hipMemcpyDeviceToHost, stream));
    RPX_HIP_TRY(hipStreamSynchronize(stream));
    if (bad) return rpx_set_error(RPX_E_INVALID, std::to_string(bad) + " gausslets have a wavelength_idx outside "
                                                 "the wavelength list");
    return RPX_OK;
}
