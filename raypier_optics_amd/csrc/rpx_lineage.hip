// Consumers of traced generations that follow parent links or vote over
// end faces, on the generations already resident in HBM:
//
//  * rpx_path_lengths -- the optical path walk of MeanOpticalPathLength
//    (results.py:112-122): for every ray of the last generation that ended
//    on the target face, the sum over its ancestry of length * Re(n),
//    accumulated from the last generation back to the first.
//  * rpx_face_histogram -- the per-generation end-face vote of
//    set_face_sequence (sources.py:138-155).
//  * rpx_propagate_neighbours -- RayCollection._eval_neighbours
//    (ctracer.pyx:1084-1106): a child's neighbour j is the child of the same
//    type (reflected / not) of its parent's neighbour j.
//
// The generation table (one entry per generation, device pointers into the
// SoA planes) is copied into the caller's workspace, so each consumer is one
// launch over all generations.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <string>
#include <vector>
#include "rpx_internal.h"
#include "rpx_host.h"
#include "rpx_gen.h"

using namespace rpx;

namespace {

struct GenDesc {
    const double *length;
    const double *n_re;
    const uint32_t *parent;
    const uint32_t *end_face;
    uint64_t n;
};

#define RPX_HIST_MAX_BINS 4094

struct LineageWs {
    GenDesc *descs;
    uint64_t *tile_status;
    GenControl *ctl;   // ctl->hits counts broken parent links
};

uint64_t lineage_bytes(int32_t n_gen, uint64_t n_last) {
    const uint64_t nb = (n_last + RPX_BLOCK - 1) / RPX_BLOCK;
    return align256(sizeof(GenDesc) * (uint64_t)std::max(n_gen, 1)) + align256(8 * std::max<uint64_t>(nb, 1)) +
           align256(sizeof(GenControl));
}

LineageWs carve_lineage(void *ws, int32_t n_gen, uint64_t n_last) {
    const uint64_t nb = (n_last + RPX_BLOCK - 1) / RPX_BLOCK;
    char *p = (char *)ws;
    LineageWs w;
    w.descs = (GenDesc *)p;
    p += align256(sizeof(GenDesc) * (uint64_t)std::max(n_gen, 1));
    w.tile_status = (uint64_t *)p;
    p += align256(8 * std::max<uint64_t>(nb, 1));
    w.ctl = (GenControl *)p;
    return w;
}

rpx_status_t upload_generations(const rpx_generation_t *gens, int32_t n_gen, GenDesc *dst, hipStream_t stream) {
    std::vector<GenDesc> h((size_t)n_gen);
    for (int32_t g = 0; g < n_gen; ++g) {
        const rpx_rays_t &r = gens[g].rays;
        const uint64_t c = r.capacity, n = gens[g].n_rays;
        if (c < n) return rpx_set_error(RPX_E_INVALID, "generation " + std::to_string(g) + ": capacity below count");
        if (n > 0 && (!r.f64 || !r.u32))
            return rpx_set_error(RPX_E_INVALID, "generation " + std::to_string(g) + " has null fields");
        h[g].length = r.f64 ? r.f64 + RPX_LENGTH * c : nullptr;
        h[g].n_re = r.f64 ? r.f64 + RPX_N_RE * c : nullptr;
        h[g].parent = r.u32 ? r.u32 + RPX_PARENT_IDX * c : nullptr;
        h[g].end_face = r.u32 ? r.u32 + RPX_END_FACE_IDX * c : nullptr;
        h[g].n = n;
    }
    RPX_HIP_TRY(hipMemcpyAsync(dst, h.data(), sizeof(GenDesc) * (size_t)n_gen, hipMemcpyHostToDevice, stream));
    // the host table must outlive the copy
    RPX_HIP_TRY(hipStreamSynchronize(stream));
    return RPX_OK;
}

}  // namespace

// --------------------------------------------------------- path lengths
// One tile of RPX_BLOCK rays of the last generation per workgroup: rays
// ending on `target` walk their parent chain (the product is rounded before
// the sum, as numpy evaluates `total += length * n.real`), and the totals are
// written in ray order, placed by the decoupled look-back.
__global__ __launch_bounds__(RPX_BLOCK) void k_path_walk(const GenDesc *G, int32_t n_gen, uint32_t target,
                                                         double *totals, uint64_t capacity, uint64_t *status,
                                                         GenControl *ctl) {
    __shared__ uint32_t lds_wave[RPX_BLOCK / 64];
    __shared__ uint32_t lds_tile;
    __shared__ uint64_t lds_excl;
    if (threadIdx.x == 0) lds_tile = atomicAdd(&ctl->tile_counter, 1u);
    __syncthreads();
    const uint32_t tile = lds_tile;
    const GenDesc last = G[n_gen - 1];
    const uint64_t i = (uint64_t)tile * RPX_BLOCK + threadIdx.x;
    const bool sel = i < last.n && last.end_face[i] == target;
    double total = 0.0;
    if (sel) {
        uint64_t idx = i;
        for (int32_t g = n_gen - 1; g >= 0; --g) {
            const double *len = G[g].length, *nre = G[g].n_re;
            const double step = len[idx] * nre[idx];
            total = total + step;
            if (g > 0) {
                const uint32_t p = G[g].parent[idx];
                if ((uint64_t)p >= G[g - 1].n) {
                    atomicAdd((unsigned long long *)&ctl->hits, 1ull);
                    break;
                }
                idx = p;
            }
        }
    }
    uint32_t block_total;
    const uint32_t local = block_scan_small(sel ? 1u : 0u, lds_wave, &block_total);
    if (threadIdx.x < 64) {
        const uint64_t excl = lookback(status, tile, 0, block_total, &ctl->error);
        if (threadIdx.x == 0) lds_excl = excl;
    }
    __syncthreads();
    const uint64_t pos = lds_excl + local;
    if (sel && pos < capacity) totals[pos] = total;
    if (threadIdx.x == 0 && (uint64_t)tile == (last.n - 1) / RPX_BLOCK) ctl->children = lds_excl + block_total;
}

extern "C" uint64_t rpx_lineage_workspace_bytes(int32_t n_generations, uint64_t n_last) {
    return lineage_bytes(n_generations, n_last);
}

extern "C" rpx_status_t rpx_path_lengths(const rpx_generation_t *generations, int32_t n_generations,
                                         uint32_t target_face, double *totals, uint64_t capacity,
                                         uint64_t *n_selected, void *workspace, uint64_t workspace_bytes,
                                         void *stream_) {
    if (!n_selected) return rpx_set_error(RPX_E_INVALID, "n_selected is null");
    *n_selected = 0;
    if (!generations || n_generations < 1) return rpx_set_error(RPX_E_INVALID, "no generations");
    const uint64_t n_last = generations[n_generations - 1].n_rays;
    if (n_last == 0) return RPX_OK;
    if (n_last > 0xFFFFFFFFull) return rpx_set_error(RPX_E_UNSUPPORTED, "more than 2^32-1 rays in one generation");
    if (capacity > 0 && !totals) return rpx_set_error(RPX_E_INVALID, "totals is null");
    if (!workspace || workspace_bytes < lineage_bytes(n_generations, n_last))
        return rpx_set_error(RPX_E_INVALID, "workspace too small");
    hipStream_t stream = (hipStream_t)stream_;
    LineageWs w = carve_lineage(workspace, n_generations, n_last);
    if (rpx_status_t st = upload_generations(generations, n_generations, w.descs, stream)) return st;
    const uint64_t nb = (n_last + RPX_BLOCK - 1) / RPX_BLOCK;
    hipLaunchKernelGGL(k_clear, dim3((unsigned)((std::max<uint64_t>(nb, RPX_SHARDS * RPX_LDS_FACES) + 255) / 256)),
                       dim3(256), 0, stream, w.tile_status, nb, w.ctl);
    RPX_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_path_walk, dim3((unsigned)nb), dim3(RPX_BLOCK), 0, stream, (const GenDesc *)w.descs,
                       n_generations, target_face, totals, capacity, w.tile_status, w.ctl);
    RPX_HIP_TRY(hipGetLastError());
    GenControl ctl;
    RPX_HIP_TRY(read_control(ctl, w.ctl, stream));
    if (ctl.error) return rpx_set_error(RPX_E_HIP, "path walk: a look-back wait exceeded its poll limit");
    *n_selected = ctl.children;
    if (ctl.hits)
        return rpx_set_error(RPX_E_INVALID, std::to_string(ctl.hits) + " selected rays have a parent_idx outside "
                                            "the previous generation");
    if (ctl.children > capacity)
        return rpx_set_error(RPX_E_CAPACITY, "totals capacity " + std::to_string(capacity) + " below " +
                                                 std::to_string(ctl.children) + " selected rays");
    return RPX_OK;
}

// ------------------------------------------------------- end-face vote
// counts[g * (n_bins + 2) + b]: rays of generation g whose end_face_idx is b
// (b < n_bins), is 0xFFFFFFFF (b = n_bins), or is any other value
// (b = n_bins + 1).  Grid: (tiles, generations); each wavefront adds runs of
// equal bins with one LDS atomic per distinct bin.
__global__ __launch_bounds__(256) void k_zero_u64(unsigned long long *p, uint64_t n) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x)
        p[k] = 0;
}

__global__ __launch_bounds__(256) void k_face_hist(const GenDesc *G, uint32_t n_bins, unsigned long long *counts) {
    __shared__ uint32_t h[RPX_HIST_MAX_BINS + 2];
    const uint32_t nb2 = n_bins + 2;
    for (uint32_t k = threadIdx.x; k < nb2; k += blockDim.x) h[k] = 0;
    __syncthreads();
    const GenDesc d = G[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    // every lane of a wavefront runs the same number of iterations
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < d.n; base += stride) {
        const uint64_t i = base + threadIdx.x;
        const bool valid = i < d.n;
        uint32_t b = 0;
        if (valid) {
            const uint32_t f = d.end_face[i];
            b = f < n_bins ? f : (f == RPX_NO_FACE ? n_bins : n_bins + 1);
        }
        uint64_t active = __ballot(valid);
        while (active) {
            const int leader = __ffsll((unsigned long long)active) - 1;
            const uint32_t lb = __shfl(b, leader);
            const uint64_t same = __ballot(valid && b == lb) & active;
            if (lane == leader) atomicAdd(&h[lb], (uint32_t)__popcll(same));
            active &= ~same;
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nb2; k += blockDim.x)
        if (h[k]) atomicAdd(&counts[(uint64_t)blockIdx.y * nb2 + k], (unsigned long long)h[k]);
}

extern "C" rpx_status_t rpx_face_histogram(const rpx_generation_t *generations, int32_t n_generations,
                                           uint32_t n_bins, uint64_t *counts, void *workspace,
                                           uint64_t workspace_bytes, void *stream_) {
    if (!generations || n_generations < 1) return rpx_set_error(RPX_E_INVALID, "no generations");
    if (n_generations > 65535) return rpx_set_error(RPX_E_UNSUPPORTED, "more than 65535 generations");
    if (n_bins > RPX_HIST_MAX_BINS)
        return rpx_set_error(RPX_E_UNSUPPORTED, "more than " + std::to_string(RPX_HIST_MAX_BINS) + " face bins");
    if (!counts) return rpx_set_error(RPX_E_INVALID, "counts is null");
    if (!workspace || workspace_bytes < lineage_bytes(n_generations, 0))
        return rpx_set_error(RPX_E_INVALID, "workspace too small");
    hipStream_t stream = (hipStream_t)stream_;
    LineageWs w = carve_lineage(workspace, n_generations, 0);
    if (rpx_status_t st = upload_generations(generations, n_generations, w.descs, stream)) return st;
    uint64_t n_max = 0;
    for (int32_t g = 0; g < n_generations; ++g) n_max = std::max(n_max, generations[g].n_rays);
    const uint64_t total = (uint64_t)n_generations * (n_bins + 2);
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)std::min<uint64_t>((total + 255) / 256, 1024)), dim3(256), 0,
                       stream, (unsigned long long *)counts, total);
    RPX_HIP_TRY(hipGetLastError());
    if (n_max > 0) {
        // about 8 rays per lane; at most 2048 workgroups per generation
        const uint64_t tiles = std::min<uint64_t>((n_max + 8 * 256 - 1) / (8 * 256), 2048);
        hipLaunchKernelGGL(k_face_hist, dim3((unsigned)tiles, (unsigned)n_generations), dim3(256), 0, stream,
                           (const GenDesc *)w.descs, n_bins, (unsigned long long *)counts);
        RPX_HIP_TRY(hipGetLastError());
    }
    RPX_HIP_TRY(hipStreamSynchronize(stream));
    return RPX_OK;
}

// ------------------------------------------------- neighbour propagation
// rmap[2 p + t] = the largest child index with parent p and type t (the
// last assignment of the reference's loop), -1 if none.
__global__ __launch_bounds__(256) void k_nb_clear(int32_t *rmap, uint64_t n, unsigned long long *err) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) rmap[k] = -1;
    if (k == 0) *err = 0;
}

__global__ __launch_bounds__(256) void k_nb_map(const uint32_t *parent, const uint32_t *type, uint64_t n_child,
                                                uint64_t n_parent, int32_t *rmap, unsigned long long *err) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_child) return;
    const uint32_t p = parent[i];
    if ((uint64_t)p >= n_parent) {
        atomicAdd(err, 1ull);
        return;
    }
    atomicMax(&rmap[2 * (uint64_t)p + (type[i] & RPX_REFL_RAY)], (int32_t)i);
}

__global__ __launch_bounds__(256) void k_nb_gather(const uint32_t *parent, const uint32_t *type, uint64_t n_child,
                                                   const int32_t *pnb, int32_t k_nb, uint64_t n_parent,
                                                   const int32_t *rmap, int32_t *nb, unsigned long long *err) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_child * (uint64_t)k_nb) return;
    const uint64_t i = t / (uint64_t)k_nb, j = t % (uint64_t)k_nb;
    const uint32_t p = parent[i];
    int32_t out = -1;
    if ((uint64_t)p < n_parent) {
        const int32_t c = pnb[(uint64_t)p * k_nb + j];
        if (c >= 0) {
            if ((uint64_t)c < n_parent) {
                out = rmap[2 * (uint64_t)c + (type[i] & RPX_REFL_RAY)];
            } else {
                atomicAdd(err, 1ull);
            }
        }
    }
    nb[t] = out;
}

extern "C" uint64_t rpx_neighbours_workspace_bytes(uint64_t n_parent) {
    return align256(8 * std::max<uint64_t>(n_parent, 1)) + 256;
}

extern "C" rpx_status_t rpx_propagate_neighbours(const rpx_rays_t *children, uint64_t n_children,
                                                 const int32_t *parent_neighbours, uint64_t n_parents,
                                                 int32_t n_neighbours, int32_t *neighbours, void *workspace,
                                                 uint64_t workspace_bytes, void *stream_) {
    if (!children) return rpx_set_error(RPX_E_INVALID, "children are null");
    if (children->capacity < n_children) return rpx_set_error(RPX_E_INVALID, "children: capacity below count");
    if (n_neighbours < 0) return rpx_set_error(RPX_E_INVALID, "negative neighbour count");
    if (n_children > 0x7FFFFFFFull || n_parents > 0x7FFFFFFFull)
        return rpx_set_error(RPX_E_UNSUPPORTED, "more than 2^31-1 rays in a generation");
    if (n_children * (uint64_t)n_neighbours > 0xFFFFFF00ull)
        return rpx_set_error(RPX_E_UNSUPPORTED, "more than 2^32-256 neighbour entries");
    if (n_children == 0 || n_neighbours == 0) return RPX_OK;
    if (!children->u32) return rpx_set_error(RPX_E_INVALID, "children have null fields");
    if (!neighbours || (n_parents > 0 && !parent_neighbours))
        return rpx_set_error(RPX_E_INVALID, "neighbour arrays are null");
    if (!workspace || workspace_bytes < rpx_neighbours_workspace_bytes(n_parents))
        return rpx_set_error(RPX_E_INVALID, "workspace too small");
    hipStream_t stream = (hipStream_t)stream_;
    int32_t *rmap = (int32_t *)workspace;
    unsigned long long *err = (unsigned long long *)((char *)workspace + align256(8 * std::max<uint64_t>(n_parents, 1)));
    const uint64_t c = children->capacity;
    const uint32_t *parent = children->u32 + RPX_PARENT_IDX * c;
    const uint32_t *type = children->u32 + RPX_RAY_TYPE_ID * c;
    const uint64_t n_map = std::max<uint64_t>(2 * n_parents, 1);
    hipLaunchKernelGGL(k_nb_clear, dim3((unsigned)((n_map + 255) / 256)), dim3(256), 0, stream, rmap, 2 * n_parents,
                       err);
    RPX_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_nb_map, dim3((unsigned)((n_children + 255) / 256)), dim3(256), 0, stream, parent, type,
                       n_children, n_parents, rmap, err);
    RPX_HIP_TRY(hipGetLastError());
    const uint64_t n_out = n_children * (uint64_t)n_neighbours;
    hipLaunchKernelGGL(k_nb_gather, dim3((unsigned)((n_out + 255) / 256)), dim3(256), 0, stream, parent, type,
                       n_children, parent_neighbours, n_neighbours, n_parents, (const int32_t *)rmap, neighbours,
                       err);
    RPX_HIP_TRY(hipGetLastError());
    unsigned long long bad = 0;
    RPX_HIP_TRY(hipMemcpyAsync(&bad, err, sizeof(bad), hipMemcpyDeviceToHost, stream));
    RPX_HIP_TRY(hipStreamSynchronize(stream));
    if (bad)
        return rpx_set_error(RPX_E_INVALID, std::to_string(bad) + " parent or neighbour indices outside the parent "
                                            "generation");
    return RPX_OK;
}
