// FP32-split ray generation (the reduced-precision trace form) and the
// conversions between the float64 and FP32-split ray layouts.
//
// The generation is the single-pass scheme of the float64 path (one launch:
// nearest face, material, decoupled look-back placement of the children),
// with the FP32-split record: the face tests run the float64 face code on the
// float32 ray promoted to float64, the materials are those of rpx_split.h.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <string>
#include "rpx_split.h"
#include "rpx_host.h"
#include "rpx_gen.h"

using namespace rpx;

struct SplitArgs {
    SceneView S;
    rpx_rays_split_t in;
    rpx_rays_split_t out;
    uint64_t n_in;
    double max_length;
    WlTable wavelengths;
    uint32_t *face_hits;
    Workspace ws;
    int32_t one_list, one_face;
    // device-resident chain (link = nullptr otherwise), as TraceArgs
    GenLink *link;
    uint64_t pool_rays;
    int32_t in_pool, store_children;
    uint64_t status_tag;
    GenControl *head_out;
    // tile order, as TraceArgs
    uint32_t tile_base = 0;
    int32_t by_launch = 0;
    uint64_t lookback_ticks = 0;
};

// The rays from element `off` of an FP32-split SoA buffer on (same stride).
RPX_DEV rpx_rays_split_t advance_split(rpx_rays_split_t r, uint64_t off) {
    r.f32 += off;
    r.f64 += off;
    r.u32 += off;
    return r;
}

// Collects the (at most two) children of one ray as a two-slot shift
// register (r0 latest), which the compiler keeps in registers.
struct PairEmitS {
    static constexpr int kCap = 2;
    RayS r0, r1;
    int n = 0;
    RPX_DEV void operator()(const RayS &x) {
        r1 = r0;
        r0 = x;
        ++n;
    }
    RPX_DEV void skip() { ++n; }
    RPX_DEV const RayS &first() const { return n > 1 ? r1 : r0; }
};

// After a generation's children are placed: stores children from index 2 on
// (a multi-order grating's third order) at pos + index, within `limit`.
struct TailEmitS {
    static constexpr int kCap = 3;
    const rpx_rays_split_t *out;
    uint64_t pos, limit;
    int n = 0;
    RPX_DEV void operator()(const RayS &r) {
        if (n >= 2 && pos + n < limit) store_ray_s(*out, pos + n, r);
        ++n;
    }
    RPX_DEV void skip() { ++n; }
};

// Rays per tile (= workgroup size) of each split variant.  The lens
// variants run 4 waves per SIMD, which two 512-lane workgroups fill exactly;
// their tiles of 512 rays halve the per-tile look-back and staging work and
// measured 4 % faster on C2 than tiles of 256.  V_FLAT (3 waves per SIMD)
// and V_ALL (2) keep 256, which their occupancy divides.
template <uint32_t F>
constexpr int split_block() {
    return (F == V_SPHERES || F == V_LENS) ? 512 : RPX_BLOCK;
}
static uint32_t split_block_of(uint32_t variant) {
    return (variant == V_SPHERES || variant == V_LENS) ? 512u : (uint32_t)RPX_BLOCK;
}

template <uint32_t F>
__global__ __launch_bounds__(split_block<F>()) __attribute__((amdgpu_waves_per_eu(F == V_ALL ? 1 : F == V_FLAT ? 3 : 4)))
void k_generation_split(SplitArgs A) {
    constexpr int B = split_block<F>();
    __shared__ SceneLds lds_scene;
    __shared__ uint32_t lds_hist[RPX_LDS_FACES];
    __shared__ uint32_t lds_wave[B / 64];
    __shared__ uint32_t lds_tile, lds_hits;
    __shared__ uint64_t lds_excl;
    // in a device-resident chain, size and pool place come from the link and
    // workgroups past the generation's tiles leave at once (as k_generation)
    uint64_t n_in = A.n_in, out_limit = A.out.capacity;
    rpx_rays_split_t in = A.in, out = A.out;
    if (A.link) {
        const GenLink L = A.link[0];
        n_in = L.count;
        if (A.in_pool) in = advance_split(A.in, L.offset);
        out = advance_split(A.out, L.out_offset);
        out_limit = A.store_children ? (A.pool_rays > L.out_offset ? A.pool_rays - L.out_offset : 0) : 0;
    }
    const uint64_t n_tiles = (n_in + B - 1) / B;
    uint32_t my_tile = 0;
    if (A.by_launch) {
        my_tile = A.tile_base + blockIdx.x;
        if (my_tile >= n_tiles) return;
    } else {
        if (blockIdx.x >= n_tiles) return;
        // the tile number is fetched first, so that its round trip overlaps
        // the scene staging
        if (threadIdx.x == 0) my_tile = atomicAdd(&A.ws.ctl->tile_counter, 1u);
    }
    // with launch-order tiles the ray's geometry is loaded before the scene
    // is staged, so that the two latencies overlap
    const float *f = in.f32;
    const uint64_t c = in.capacity;
    F3 o32 = f3(0, 0, 0), d32 = f3(0, 0, 0);
    if (A.by_launch) {
        const uint64_t i0 = (uint64_t)my_tile * B + threadIdx.x;
        if (i0 < n_in) {
            o32 = f3(f[RPX_S_OX * c + i0], f[RPX_S_OY * c + i0], f[RPX_S_OZ * c + i0]);
            d32 = f3(f[RPX_S_DX * c + i0], f[RPX_S_DY * c + i0], f[RPX_S_DZ * c + i0]);
        }
    }
    const SceneView S = stage_scene(A.S, lds_scene);
    for (int k = threadIdx.x; k < RPX_LDS_FACES; k += blockDim.x) lds_hist[k] = 0;
    if (threadIdx.x == 0) {
        lds_tile = my_tile;
        lds_hits = 0;
    }
    __syncthreads();
    const uint32_t tile = lds_tile;
    const uint64_t i = (uint64_t)tile * B + threadIdx.x;
    PairEmitS ce;
    bool hit = false;
    int hit_rec = -1, hit_list = -1;
    double hit_length = 0;
    if (i < n_in) {
        if (!A.by_launch) {
            o32 = f3(f[RPX_S_OX * c + i], f[RPX_S_OY * c + i], f[RPX_S_OZ * c + i]);
            d32 = f3(f[RPX_S_DX * c + i], f[RPX_S_DY * c + i], f[RPX_S_DZ * c + i]);
        }
        const V3 origin = v3_of(o32);
        const V3 dir = v3_of(d32);
        double length = A.max_length;
        uint32_t end_face = RPX_NO_FACE;
        const V3 point = add(origin, mul(dir, A.max_length));
        const int l0 = A.one_list < 0 ? 0 : A.one_list;
        const int l1 = A.one_list < 0 ? S.n_face_lists : A.one_list + 1;
        for (int l = l0; l < l1; ++l) {
            const rpx_face_list_t &fl = S.face_lists[l];
            const V3 p1 = transform(fl.inv_trans, origin);
            const V3 p2 = transform(fl.inv_trans, point);
            const double slen = sep(p1, p2), zrate = (p2.z - p1.z) / slen;
            const int k0 = A.one_list < 0 ? fl.first_face : fl.first_face + A.one_face;
            const int k1 = A.one_list < 0 ? fl.first_face + fl.n_faces : k0 + 1;
            for (int k = k0; k < k1; ++k) {
                auto test = [&](const rpx_face_t &fc) {
                    if (face_out_of_reach(fc, p1.z, zrate * length, slen)) return;
                    const double dist = face_intersect<F>(S, fc, p1, p2, 1);
                    if (fc.tolerance < dist && dist < length) {
                        length = dist;
                        end_face = (uint32_t)fc.global_idx;
                        hit_rec = k;
                        hit_list = l;
                    }
                };
                // k is the same in every lane: the lens and sphere variants
                // read the face record with scalar loads (SGPRs, not VGPRs,
                // hold its parameters through the root solve), which lets
                // them run 4 waves per SIMD; the all-features variant reads
                // the LDS copy
                if constexpr (F == V_ALL) test(S.faces[k]);
                else test(face_uniform(A.S.faces, k));
            }
        }
        in.f32[RPX_S_LENGTH * c + i] = (float)length;
        in.u32[RPX_END_FACE_IDX * c + i] = end_face;
        hit_length = length;
        if (hit_rec >= 0) {
            atomicAdd(&lds_hist[hit_rec], 1u);
            const rpx_face_t &fc = S.faces[hit_rec];
            if (fc.material >= 0 && material_emits(S.materials[fc.material])) {
                hit = true;   // counted: the rest of its record is read
                const rpx_face_list_t &fl = S.face_lists[hit_list];
                RayS r = load_ray_s(in, i);
                r.length = (float)length;
                r.end_face_idx = end_face;
                const V3 hp = add(origin, mul(dir, length));
                const Orientation o = compute_orientation<F>(S, fl, fc, hp);
                eval_children_split<F>(S.materials[fc.material], r, length, (uint32_t)i, hp, o, A.wavelengths, ce);
            }
        }
    }
    // a multi-order grating (all-features variant only) can have a third child
    const uint32_t count = (uint32_t)min(ce.n, (F & M_GRATING) ? 3 : 2);
    uint32_t block_total;
    const uint32_t local = block_scan_small(count, lds_wave, &block_total);
    {
        const uint64_t mh = __ballot(hit);
        if ((threadIdx.x & 63) == 0) atomicAdd(&lds_hits, (uint32_t)__popcll(mh));
    }
    if (threadIdx.x < 64) {
        const uint64_t excl = lookback_tagged(A.ws.tile_status, tile, 0, block_total, &A.ws.ctl->error, A.status_tag,
                                              A.lookback_ticks);
        if (threadIdx.x == 0) lds_excl = excl;
    }
    __syncthreads();
    const uint64_t pos = lds_excl + local;
    if (count > 0 && pos < out_limit) store_ray_s(out, pos, ce.first());
    if (count > 1 && pos + 1 < out_limit) store_ray_s(out, pos + 1, ce.r0);
    if constexpr ((F & M_GRATING) != 0) {
        if (count > 2 && pos + 2 < out_limit) {
            // the third child is built again now that its position is known
            const rpx_face_t &fc = S.faces[hit_rec];
            RayS r = load_ray_s(in, i);
            r.length = (float)hit_length;
            const V3 hp = add(v3_of(r.origin), mul(v3_of(r.direction), hit_length));
            const Orientation o = compute_orientation<F>(S, S.face_lists[hit_list], fc, hp);
            TailEmitS te{&out, pos, out_limit};
            grating_children_split(S.materials[fc.material], r, hit_length, (uint32_t)i, hp, o, A.wavelengths, te);
        }
    }
    if (threadIdx.x == 0) {
        // a chain's hit total is the sum of its face totals (k_chain_fold);
        // one word per generation would take every tile's atomic in one line
        if (!A.link) add_hits(A.ws.ctl, tile, lds_hits);
        if ((uint64_t)tile == (n_in - 1) / B) {
            const uint64_t kids = lds_excl + block_total;
            A.ws.ctl->children = kids;
            if (A.link) {
                const uint64_t at = A.link[0].out_offset;
                A.link[1].count = kids;
                A.link[1].offset = at;
                A.link[1].out_offset = at + kids;
                A.head_out->children = kids;   // host-mapped: read once the generation has ended
            }
        }
    }
    if (A.face_hits || A.link) {
        uint32_t *shard = A.ws.ctl->face_shards[tile % RPX_SHARDS];
        for (int k = threadIdx.x; k < S.n_list_faces; k += blockDim.x) {
            uint32_t h = lds_hist[k];
            if (h) atomicAdd(&shard[k], h);
        }
    }
}

__global__ __launch_bounds__(256) void k_to_split(rpx_rays_t in, uint64_t n, rpx_rays_split_t out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    static_assert(RPX_SPLIT_F32_FIELDS == RPX_LENGTH + 1, "split float fields mirror the float64 ones");
    for (int k = 0; k < RPX_SPLIT_F32_FIELDS; ++k) out.f32[k * out.capacity + i] = (float)in.f64[k * in.capacity + i];
    out.f64[RPX_S_PHASE * out.capacity + i] = in.f64[RPX_PHASE * in.capacity + i];
    out.f64[RPX_S_ACCUM * out.capacity + i] = in.f64[RPX_ACCUM * in.capacity + i];
    for (int k = 0; k < RPX_N_U32_FIELDS; ++k) out.u32[k * out.capacity + i] = in.u32[k * in.capacity + i];
}

__global__ __launch_bounds__(256) void k_from_split(rpx_rays_split_t in, uint64_t n, rpx_rays_t out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < RPX_SPLIT_F32_FIELDS; ++k) out.f64[k * out.capacity + i] = (double)in.f32[k * in.capacity + i];
    out.f64[RPX_PHASE * out.capacity + i] = in.f64[RPX_S_PHASE * in.capacity + i];
    out.f64[RPX_ACCUM * out.capacity + i] = in.f64[RPX_S_ACCUM * in.capacity + i];
    for (int k = 0; k < RPX_N_U32_FIELDS; ++k) out.u32[k * out.capacity + i] = in.u32[k * in.capacity + i];
}

static rpx_status_t check_split(const rpx_rays_split_t *r, uint64_t n, const char *what) {
    if (!r) return rpx_set_error(RPX_E_INVALID, std::string(what) + " is null");
    if (r->capacity < n) return rpx_set_error(RPX_E_INVALID, std::string(what) + " capacity below count");
    if (n > 0 && (!r->f32 || !r->f64 || !r->u32))
        return rpx_set_error(RPX_E_INVALID, std::string(what) + " has null fields");
    return RPX_OK;
}

static rpx_status_t check_f64(const rpx_rays_t *r, uint64_t n, const char *what) {
    if (!r) return rpx_set_error(RPX_E_INVALID, std::string(what) + " is null");
    if (r->capacity < n) return rpx_set_error(RPX_E_INVALID, std::string(what) + " capacity below count");
    if (n > 0 && (!r->f64 || !r->u32)) return rpx_set_error(RPX_E_INVALID, std::string(what) + " has null fields");
    return RPX_OK;
}

extern "C" rpx_status_t rpx_rays_to_split(const rpx_rays_t *in, uint64_t n, rpx_rays_split_t *out, void *stream) {
    rpx_status_t st = check_f64(in, n, "input rays");
    if (st) return st;
    if ((st = check_split(out, n, "output rays"))) return st;
    if (n == 0) return RPX_OK;
    hipLaunchKernelGGL(k_to_split, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, *in, n,
                       *out);
    RPX_HIP_TRY(hipGetLastError());
    return RPX_OK;
}

extern "C" rpx_status_t rpx_rays_from_split(const rpx_rays_split_t *in, uint64_t n, rpx_rays_t *out,
                                            void *stream) {
    rpx_status_t st = check_split(in, n, "input rays");
    if (st) return st;
    if ((st = check_f64(out, n, "output rays"))) return st;
    if (n == 0) return RPX_OK;
    hipLaunchKernelGGL(k_from_split, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, *in,
                       n, *out);
    RPX_HIP_TRY(hipGetLastError());
    return RPX_OK;
}

// The scene's variant, or V_FLAT for a scene that needs V_ALL only for its
// distortions and gratings: at 3 waves per SIMD (with some spilling) it runs
// a Michelson of a beamsplitter cube, a plane and a Zernike mirror and a
// grating (1,048,576 rays to depth 12) 5 % faster than V_ALL at 2.
static uint32_t split_variant(const rpx_scene_t *s) {
    const uint32_t v = scene_variant(s);
    return (v == V_ALL && (s->features & ~(uint32_t)V_FLAT) == 0) ? (uint32_t)V_FLAT : v;
}

static void launch_split(uint32_t variant, dim3 grid, hipStream_t stream, const SplitArgs &A) {
    const dim3 block(split_block_of(variant));
    if (variant == V_SPHERES)
        hipLaunchKernelGGL((k_generation_split<V_SPHERES>), grid, block, 0, stream, A);
    else if (variant == V_FLAT)
        hipLaunchKernelGGL((k_generation_split<V_FLAT>), grid, block, 0, stream, A);
    else if (variant == V_LENS)
        hipLaunchKernelGGL((k_generation_split<V_LENS>), grid, block, 0, stream, A);
    else
        hipLaunchKernelGGL((k_generation_split<V_ALL>), grid, block, 0, stream, A);
}

extern "C" rpx_status_t rpx_trace_segment_split(const rpx_scene_t *scene, int32_t face_list, int32_t face_idx,
                                                const rpx_rays_split_t *in, uint64_t n_in, float max_length,
                                                const double *wavelengths, uint64_t n_wavelengths,
                                                rpx_rays_split_t *out, uint64_t *n_out, uint32_t *face_hits,
                                                void *workspace, uint64_t workspace_bytes, void *stream_) {
    if (!n_out) return rpx_set_error(RPX_E_INVALID, "n_out is null");
    *n_out = 0;
    rpx_status_t st = rpx_check_trace_scene(scene);
    if (st) return st;
    st = check_split(in, n_in, "input rays");
    if (st) return st;
    if (!out) return rpx_set_error(RPX_E_INVALID, "output rays are null");
    if (out->capacity > 0 && (!out->f32 || !out->f64 || !out->u32))
        return rpx_set_error(RPX_E_INVALID, "output rays have null fields");
    if (n_in == 0) return RPX_OK;
    if (n_in > 0xFFFFFFFFull) return rpx_set_error(RPX_E_UNSUPPORTED, "more than 2^32-1 rays in one generation");
    if (!wavelengths && n_wavelengths) return rpx_set_error(RPX_E_INVALID, "wavelength table is null");
    if (!workspace || workspace_bytes < rpx_trace_workspace_bytes(n_in))
        return rpx_set_error(RPX_E_INVALID, "workspace too small");
    if (face_list >= 0 && (face_list >= scene->view.n_face_lists || face_idx < 0 ||
                           face_idx >= scene->list_sizes[face_list]))
        return rpx_set_error(RPX_E_INVALID, "face list / face index out of range");
    hipStream_t stream = (hipStream_t)stream_;
    SplitArgs A;
    A.S = scene->view;
    A.in = *in;
    A.out = *out;
    A.n_in = n_in;
    A.max_length = (double)max_length;
    A.wavelengths.wl = wavelengths;
    A.wavelengths.n = wavelengths ? n_wavelengths : 0;
    A.face_hits = face_hits;
    A.ws = carve(workspace, n_in);
    A.one_list = face_list < 0 ? -1 : face_list;
    A.one_face = face_list < 0 ? -1 : face_idx;
    A.link = nullptr;
    A.pool_rays = 0;
    A.in_pool = 0;
    A.store_children = 1;
    A.status_tag = 0;
    A.head_out = nullptr;
    const uint32_t block = split_block_of(split_variant(scene));
    const uint64_t nb = (n_in + block - 1) / block;
    static thread_local EventSet ev;
    const bool timing = rpx_timing_is_on() && ev.ok;
    if (timing) RPX_HIP_TRY(hipEventRecord(ev.e[0], stream));
    hipLaunchKernelGGL(k_clear, dim3((unsigned)((std::max<uint64_t>(nb, RPX_SHARDS * RPX_LDS_FACES) + 255) / 256)),
                           dim3(256), 0, stream, A.ws.tile_status, nb,
                       A.ws.ctl);
    RPX_HIP_TRY(hipGetLastError());
    if (timing) RPX_HIP_TRY(hipEventRecord(ev.e[1], stream));
    launch_split(split_variant(scene), dim3((unsigned)nb), stream, A);
    RPX_HIP_TRY(hipGetLastError());
    if (timing) RPX_HIP_TRY(hipEventRecord(ev.e[2], stream));
    GenControl ctl;
    RPX_HIP_TRY(finish_generation(ctl, A.ws.ctl, scene->view.faces, scene->view.n_list_faces, face_hits,
                                  stream));
    if (ctl.error) return rpx_set_error(RPX_E_HIP, "generation scan: a look-back wait exceeded its poll limit");
    ctl.hits = 0;
    for (int s2 = 0; s2 < RPX_SHARDS; ++s2) ctl.hits += ctl.hit_shards[s2];
    *n_out = ctl.children;
    if (timing) {
        float t_clear = 0, t_gen = 0;
        RPX_HIP_TRY(hipEventElapsedTime(&t_clear, ev.e[0], ev.e[1]));
        RPX_HIP_TRY(hipEventElapsedTime(&t_gen, ev.e[1], ev.e[2]));
        rpx_timing_add(RPX_TIMING_CLEAR, t_clear, (double)nb * 8.0);
        // algorithmic bytes, FP32-split record: geometry 24 B read and
        // length + end face 8 B written per input ray, the other 64 B of each
        // ray that hit, 112 B per child
        rpx_timing_add(RPX_TIMING_SPLIT, t_gen,
                       (double)n_in * (24.0 + 8.0) + (double)ctl.hits * 64.0 + (double)ctl.children * 112.0);
    }
    if (ctl.children > out->capacity)
        return rpx_set_error(RPX_E_CAPACITY, "output capacity " + std::to_string(out->capacity) + " below " +
                                                 std::to_string(ctl.children) + " children");
    return RPX_OK;
}

// rpx_trace_rays_split: the FP32-split generations of trace_rays as one
// device-resident chain (run_chain in rpx_gen.h).
extern "C" rpx_status_t rpx_trace_rays_split(const rpx_scene_t *scene, const rpx_rays_split_t *in, uint64_t n_in,
                                             float max_length, const double *wavelengths, uint64_t n_wavelengths,
                                             int32_t max_generations, rpx_rays_split_t *pool, uint64_t *counts,
                                             uint64_t *offsets, int32_t *n_traced, uint32_t *face_hits,
                                             void *workspace, uint64_t workspace_bytes, void *stream_) {
    if (!n_traced || !counts || !offsets)
        return rpx_set_error(RPX_E_INVALID, "counts / offsets / n_traced are null");
    *n_traced = 0;
    counts[0] = n_in;
    offsets[0] = 0;
    if (rpx_status_t st = rpx_check_trace_scene(scene)) return st;
    if (rpx_status_t st = check_split(in, n_in, "input rays")) return st;
    if (!pool || (pool->capacity > 0 && (!pool->f32 || !pool->f64 || !pool->u32)))
        return rpx_set_error(RPX_E_INVALID, "pool");
    if (max_generations < 0) return rpx_set_error(RPX_E_INVALID, "max_generations is negative");
    if (n_in == 0 || max_generations == 0) return RPX_OK;
    if (!wavelengths && n_wavelengths) return rpx_set_error(RPX_E_INVALID, "wavelength table is null");
    if (!face_hits && scene->view.n_list_faces > 0) return rpx_set_error(RPX_E_INVALID, "face_hits is null");
    ChainSpec spec;
    spec.n_in = n_in;
    spec.pool_rays = pool->capacity;
    spec.max_generations = max_generations;
    spec.max_children = (uint64_t)std::max(scene->view.max_children, 0);
    spec.timing_slot = RPX_TIMING_SPLIT;
    spec.bytes_in = 24.0 + 8.0;    // FP32-split geometry read, length + end face written
    spec.bytes_hit = 64.0;
    spec.bytes_child = 112.0;
    if (rpx_status_t st = check_chain(spec, scene, workspace, workspace_bytes)) return st;
    const uint32_t variant = split_variant(scene);
    spec.block = split_block_of(variant);
    SplitArgs A;
    A.S = scene->view;
    A.out = *pool;
    A.n_in = 0;
    A.max_length = (double)max_length;
    A.wavelengths.wl = wavelengths;
    A.wavelengths.n = wavelengths ? n_wavelengths : 0;
    A.face_hits = nullptr;
    A.ws.tags = nullptr;
    A.ws.block_totals = nullptr;
    A.ws.block_parents = nullptr;
    A.ws.block_offsets = nullptr;
    A.one_list = -1;
    A.one_face = -1;
    A.pool_rays = pool->capacity;
    auto launch = [&](const ChainGen &cg, uint64_t grid, hipStream_t stream) {
        A.link = cg.link;
        A.ws.ctl = cg.ctl;
        A.ws.tile_status = cg.tile_status;
        A.status_tag = cg.tag;
        A.head_out = cg.head_out;
        A.in_pool = cg.in_pool;
        A.in = cg.in_pool ? *pool : *in;
        A.store_children = cg.store_children;
        A.tile_base = cg.tile_base;
        A.by_launch = cg.by_launch;
        A.lookback_ticks = cg.lookback_ticks;
        launch_split(variant, dim3((unsigned)grid), stream, A);
        return hipGetLastError();
    };
    return run_chain(spec, scene, workspace, launch, counts, offsets, n_traced, face_hits, (hipStream_t)stream_);
}
