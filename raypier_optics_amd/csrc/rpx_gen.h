// Generation machinery shared by the FP64 and FP32-split trace kernels:
// block size and LDS limits, the per-generation workspace, scene staging in
// LDS, ballot scans and the decoupled look-back that places children.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include <hip/hip_runtime.h>
#include <stddef.h>
#include "rpx_internal.h"
#include "rpx_materials.h"
#include "rpx_host.h"

namespace rpx {

#define RPX_LDS_FACES 64
#define RPX_LDS_LISTS 16
#define RPX_BLOCK 256

// ------------------------------------------------------------ workspace
// Generation control words (device), cleared by k_clear before each
// single-pass generation and read back by the host afterwards.
// Hit counters are split into RPX_SHARDS copies, chosen by tile number, so
// that the workgroups of a generation do not all add to the same address.
#define RPX_SHARDS 32
struct GenControl {
    uint64_t children;       // children written (grand total)
    uint64_t parents;        // rays with at least one child (gausslet passes)
    uint64_t hits;           // rays whose nearest-face search found a face (gausslet passes)
    uint64_t error;          // nonzero: a look-back wait ran out of polls
    uint32_t tile_counter;   // next tile id handed to a starting workgroup
    uint32_t wg_done;        // device-resident chain: workgroups that have finished their tile
    uint32_t pad[6];
    uint32_t hit_shards[RPX_SHARDS];                  // single pass: rays that hit, per shard
    uint32_t parent_shards[RPX_SHARDS];               // single-pass gausslets: rays with children, per shard
    uint32_t face_totals[RPX_LDS_FACES];              // device-resident chain: hits per face record
    uint32_t face_shards[RPX_SHARDS][RPX_LDS_FACES];  // single pass: hits per face record, per shard
};
// bytes of GenControl the host reads back after a single-pass generation
#define RPX_CTL_HEAD (offsetof(GenControl, face_shards))

struct Workspace {
    uint32_t *tags;          // per ray: child count (bits 0..7) | (face record + 1) << 8
    uint32_t *block_totals;  // per block: children
    uint32_t *block_parents; // per block: rays with at least one child
    uint64_t *block_offsets; // per block, exclusive
    uint64_t *tile_status;   // per tile: look-back flag (bits 62..63) | value
    GenControl *ctl;
};

static uint64_t align256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

// One generation of a device-resident chain: its ray count, its first ray's
// position in the pool, and where its children go (offset + count for a
// generation in the pool; 0 for the chain's first, caller-owned generation).
struct GenLink {
    uint64_t count, offset, out_offset, pad;
};

// The rays from element `off` of a SoA buffer on (same field stride).
RPX_DEV rpx_rays_t advance_rays(rpx_rays_t r, uint64_t off) {
    r.f64 += off;
    r.u32 += off;
    if (r.para) r.para += off;
    return r;
}

static Workspace carve(void *ws, uint64_t n_in) {
    uint64_t nb = (n_in + RPX_BLOCK - 1) / RPX_BLOCK;
    char *p = (char *)ws;
    Workspace w;
    w.tags = (uint32_t *)p;
    p += align256(4 * n_in);
    w.block_totals = (uint32_t *)p;
    p += align256(4 * nb);
    w.block_parents = (uint32_t *)p;
    p += align256(4 * nb);
    w.block_offsets = (uint64_t *)p;
    p += align256(8 * nb);
    w.tile_status = (uint64_t *)p;
    p += align256(8 * nb);
    w.ctl = (GenControl *)p;
    return w;
}

// Stage face lists and faces in LDS; hand back a view that points there.
// LDS copy of a scene's small tables.  Faces and face lists always fit
// (validated at scene creation); materials and shape programs are staged when
// they fit and otherwise read from global memory.
#define RPX_LDS_MATERIALS 8
#define RPX_LDS_SHAPE_OPS 32
struct SceneLds {
    rpx_face_t faces[RPX_LDS_FACES];
    rpx_face_list_t lists[RPX_LDS_LISTS];
    rpx_material_t materials[RPX_LDS_MATERIALS];
    rpx_shape_op_t shape_ops[RPX_LDS_SHAPE_OPS];
};

template <class T>
static __device__ __forceinline__ void lds_copy(T *dst, const T *src, int n) {
    static_assert(sizeof(T) % 8 == 0, "tables are copied as doubles");
    const int nw = n * (int)(sizeof(T) / 8);
    const double *s = (const double *)src;
    double *d = (double *)dst;
    for (int k = threadIdx.x; k < nw; k += blockDim.x) d[k] = s[k];
}

// Stage the scene in LDS; hand back a view that points there.  Ends with a
// workgroup barrier.
static __device__ __forceinline__ SceneView stage_scene(const SceneView &S, SceneLds &L) {
    lds_copy(L.faces, S.faces, S.n_faces);
    lds_copy(L.lists, S.face_lists, S.n_face_lists);
    const bool mats = S.n_materials <= RPX_LDS_MATERIALS, ops = S.n_shape_ops <= RPX_LDS_SHAPE_OPS;
    if (mats) lds_copy(L.materials, S.materials, S.n_materials);
    if (ops) lds_copy(L.shape_ops, S.shape_ops, S.n_shape_ops);
    __syncthreads();
    SceneView v = S;
    v.faces = L.faces;
    v.face_lists = L.lists;
    if (mats) v.materials = L.materials;
    if (ops) v.shape_ops = L.shape_ops;
    return v;
}

// Block-wide exclusive scan of small per-lane counts (0..7) with wavefront
// ballots: three bit-planes, popcount below the lane, wave totals in LDS.
static __device__ __forceinline__ uint32_t block_scan_small(uint32_t c, uint32_t *lds_wave, uint32_t *block_total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        uint64_t m = __ballot((c >> b) & 1u);
        pre += (uint32_t)__popcll(m & lt) << b;
        tot += (uint32_t)__popcll(m) << b;
    }
    if (lane == 0) lds_wave[wave] = tot;
    __syncthreads();
    uint32_t base = 0, all = 0;
    for (int w = 0; w < nwave; ++w) {
        uint32_t v = lds_wave[w];
        if (w < wave) base += v;
        all += v;
    }
    *block_total = all;
    return base + pre;
}

// --------------------------------------------------- single-pass generation
// One launch per ray generation: intersect, evaluate the material once, and
// place the children with a decoupled look-back scan over tiles of RPX_BLOCK
// rays.  Tiles are handed out by an atomic counter in launch order, so every
// tile a workgroup waits for belongs to a workgroup that is already running;
// the wait is also bounded by a poll limit (error flag instead of a hang).

#define RPX_LOOKBACK_POLLS (1u << 22)
// A status word: flag (bits 62-63) | tag (bits 48-61) | value (bits 0-47).
// A word whose tag is not the scan's own reads as unpublished, so the
// generations of a device-resident chain share one status array without
// clearing it between generations (each has its own tag).
static constexpr uint64_t ST_AGG = 1ull << 62;     // tile aggregate published
static constexpr uint64_t ST_PREFIX = 2ull << 62;  // inclusive prefix published
static constexpr uint64_t ST_VALUE = (1ull << 48) - 1;
static constexpr int ST_TAG_SHIFT = 48;
static constexpr uint64_t ST_TAG_MASK = (1ull << 14) - 1;

static __global__ __launch_bounds__(256) void k_clear(uint64_t *tile_status, uint64_t nb, GenControl *ctl) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nb) tile_status[k] = 0;
    if (k < RPX_SHARDS * RPX_LDS_FACES) (&ctl->face_shards[0][0])[k] = 0;
    if (k < RPX_SHARDS) {
        ctl->hit_shards[k] = 0;
        ctl->parent_shards[k] = 0;
    }
    if (k < RPX_LDS_FACES) ctl->face_totals[k] = 0;
    if (k == 0) {
        ctl->children = 0;
        ctl->parents = 0;
        ctl->hits = 0;
        ctl->error = 0;
        ctl->tile_counter = 0;
        ctl->wg_done = 0;
    }
}

// Ends a single-pass generation.  Adds its sharded per-face-record hit
// counts to face_hits (indexed by global face index) when face_hits is set,
// and copies the control head to head_out (host memory mapped for the
// device) when that is set.  One workgroup.
static __global__ __launch_bounds__(64) void k_fold_hits(const GenControl *ctl, const rpx_face_t *faces,
                                                          int n_list_faces, uint32_t *face_hits,
                                                          GenControl *head_out) {
    const int k = threadIdx.x;
    if (head_out) {
        const uint32_t *src = (const uint32_t *)ctl;
        uint32_t *dst = (uint32_t *)head_out;
        for (int i = k; i < (int)(RPX_CTL_HEAD / 4); i += 64) dst[i] = src[i];
    }
    if (!face_hits || k >= n_list_faces) return;
    uint32_t h = 0;
    for (int s = 0; s < RPX_SHARDS; ++s) h += ctl->face_shards[s][k];
    if (h) atomicAdd(&face_hits[faces[k].global_idx], h);
}

// Single-pass epilogue of workgroup 0-thread work: hit counts into the shards.
static __device__ __forceinline__ void add_hits(GenControl *ctl, uint64_t tile, uint32_t hits) {
    if (hits) atomicAdd(&ctl->hit_shards[tile % RPX_SHARDS], hits);
}

static __device__ __forceinline__ uint64_t status_load(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ void status_store(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Run by one whole wavefront.  Publishes the tile's aggregate, sums the
// predecessors' aggregates 64 tiles at a time until an inclusive prefix is
// found, publishes this tile's inclusive prefix and returns the exclusive one.
// The scan starts at tile `first` (tiles before it belong to another scan)
// and reads only words carrying its `tag`; a wait that runs out of polls, or
// (max_ticks > 0) lasts longer than max_ticks ticks of the 100 MHz-class
// real-time counter, adds 1 to *error and ends the wait with the part of the
// prefix seen so far (never more than the true prefix, so every position
// derived from it stays within the true bounds).  max_ticks ==
// RPX_LOOKBACK_EXPIRED ends the first wait at once (a test hook).
#define RPX_LOOKBACK_EXPIRED 1ull
static __device__ uint64_t lookback_tagged(uint64_t *status, uint64_t tile, uint64_t first, uint64_t agg,
                                           uint64_t *error, uint64_t tag, uint64_t max_ticks = 0) {
    const int lane = threadIdx.x & 63;
    const uint64_t tagbits = (tag & ST_TAG_MASK) << ST_TAG_SHIFT;
    if (tile == first) {
        if (lane == 0) status_store(&status[tile], ST_PREFIX | tagbits | agg);
        return 0;
    }
    if (lane == 0) status_store(&status[tile], ST_AGG | tagbits | agg);
    uint64_t excl = 0;
    int64_t end = (int64_t)tile - 1;
    uint32_t polls = 0;
    const uint64_t t0 = max_ticks ? __builtin_amdgcn_s_memrealtime() : 0;
    for (;;) {
        const int64_t idx = end - lane;
        uint64_t st = (idx >= (int64_t)first) ? status_load(&status[idx]) : (ST_PREFIX | tagbits);
        if (((st >> ST_TAG_SHIFT) & ST_TAG_MASK) != (tag & ST_TAG_MASK)) st = 0;   // an older scan's word
        const uint64_t flag = st >> 62;
        const uint64_t pm = __ballot(flag == 2);
        const int first_prefix = pm ? (__ffsll((unsigned long long)pm) - 1) : 64;
        const uint64_t window = (first_prefix >= 63) ? ~0ull : ((2ull << first_prefix) - 1);
        if (__ballot(flag == 0) & window) {
            bool expired = ++polls > RPX_LOOKBACK_POLLS;
            if (max_ticks)
                expired = expired || max_ticks == RPX_LOOKBACK_EXPIRED ||
                          __builtin_amdgcn_s_memrealtime() - t0 > max_ticks;
            if (expired) {
                if (lane == 0) atomicAdd((unsigned long long *)error, 1ull);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
            continue;
        }
        uint64_t v = (lane <= first_prefix) ? (st & ST_VALUE) : 0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        excl += v;
        if (first_prefix < 64) break;
        end -= 64;
    }
    if (lane == 0) status_store(&status[tile], ST_PREFIX | tagbits | (excl + agg));
    return excl;
}

// The untagged scan of a status array cleared before each use.
static __device__ uint64_t lookback(uint64_t *status, uint64_t tile, uint64_t first, uint64_t agg,
                                    uint64_t *error) {
    return lookback_tagged(status, tile, first, agg, error, 0);
}

// A pinned per-thread block, mapped for every device, that control heads are
// read back into (a pageable destination would cost a staged copy); held
// until the process ends.  Both pointers are null if it cannot be allocated.
struct PinnedControl {
    GenControl *host = nullptr, *dev = nullptr;
};
static inline const PinnedControl &pinned_control() {
    thread_local PinnedControl p;
    thread_local bool tried = false;
    if (!tried) {
        tried = true;
        GenControl *h = nullptr;
        void *d = nullptr;
        if (hipHostMalloc((void **)&h, sizeof(GenControl), hipHostMallocMapped | hipHostMallocPortable) ==
                hipSuccess &&
            hipHostGetDevicePointer(&d, h, 0) == hipSuccess) {
            p.host = h;
            p.dev = (GenControl *)d;
        }
    }
    return p;
}

// Reads the head of a control block back to the host and waits for it.
static inline hipError_t read_control(GenControl &out, const GenControl *d_ctl, hipStream_t stream) {
    GenControl *dst = pinned_control().host ? pinned_control().host : &out;
    hipError_t e = hipMemcpyAsync(dst, d_ctl, RPX_CTL_HEAD, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e == hipSuccess && dst != &out) memcpy(&out, dst, RPX_CTL_HEAD);
    return e;
}

// Ends a single-pass generation: one k_fold_hits launch folds the face hits
// (when face_hits is set) and writes the control head into the pinned block,
// which the host reads once the stream has drained.
static inline hipError_t finish_generation(GenControl &out, const GenControl *d_ctl, const rpx_face_t *faces,
                                           int n_list_faces, uint32_t *face_hits, hipStream_t stream) {
    const PinnedControl &p = pinned_control();
    if (p.host || face_hits) {
        hipLaunchKernelGGL(k_fold_hits, dim3(1), dim3(64), 0, stream, d_ctl, faces, n_list_faces, face_hits, p.dev);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (!p.host) return read_control(out, d_ctl, stream);
    hipError_t e = hipStreamSynchronize(stream);
    if (e == hipSuccess) memcpy(&out, p.host, RPX_CTL_HEAD);
    return e;
}

// ------------------------------------------------ device-resident chain
// trace_rays' generations launched back to back on one stream, for the
// float64 and the FP32-split generation kernels alike.  Generation g reads
// its size and pool position from link[g] (written by the last tile of
// generation g-1) and writes its children's to link[g+1]; look-back status
// words carry the generation as their tag, so one status array serves the
// chain without clears; each generation has a control block of its own, and
// its last tile writes its child count to a host-mapped head.  The host
// learns generation g's size once generation g-1 has ended, which it waits
// for while generation g runs (one generation late, so the GPU never idles
// on the host), sizes generation g's grid from the last observed
// children-per-parent ratio (workgroups past the tiles leave at once; tiles
// past the grid get a second launch before the next generation), and only
// launches a generation whose children the pool is sure to hold -- else it
// stops with RPX_E_CAPACITY, generation g stored and untraced, for the
// caller to continue with a fresh pool.  Hit totals are folded once, after
// the last generation.

struct ChainWs {
    uint64_t *tile_status;
    GenControl *ctl;
    GenLink *link;
};
static inline uint64_t chain_tiles(uint64_t max_rays) { return (max_rays + RPX_BLOCK - 1) / RPX_BLOCK; }
static inline uint64_t chain_workspace_bytes(uint64_t max_rays, int32_t max_generations) {
    if (max_generations < 1) max_generations = 1;
    return align256(8 * std::max<uint64_t>(chain_tiles(max_rays), 1)) +
           align256(sizeof(GenControl) * (uint64_t)(max_generations + 1)) +
           align256(sizeof(GenLink) * (uint64_t)(max_generations + 1));
}
static inline ChainWs carve_chain(void *ws, uint64_t max_rays, int32_t max_generations) {
    char *p = (char *)ws;
    ChainWs w;
    w.tile_status = (uint64_t *)p;
    p += align256(8 * std::max<uint64_t>(chain_tiles(max_rays), 1));
    w.ctl = (GenControl *)p;
    p += align256(sizeof(GenControl) * (uint64_t)(max_generations + 1));
    w.link = (GenLink *)p;
    return w;
}

// Clears a chain's workspace and host heads; link[0] describes the input.
static __global__ void k_chain_init(ChainWs w, uint64_t n_tiles, int32_t n_gen, uint64_t n_in, GenControl *heads) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n_tiles) w.tile_status[k] = 0;
    const uint64_t words = sizeof(GenControl) / 4 * (uint64_t)(n_gen + 1);
    if (k < words) {
        ((uint32_t *)w.ctl)[k] = 0;
        if (heads) ((uint32_t *)heads)[k] = 0;
    }
    if (k <= (uint64_t)n_gen) {
        w.link[k].count = k == 0 ? n_in : 0;
        w.link[k].offset = 0;
        w.link[k].out_offset = 0;
        w.link[k].pad = 0;
    }
}

// Per-face and ray hit totals and the look-back error count of every traced
// generation of a chain into its host head (the generations themselves write
// only their child count there).  One workgroup per generation.
static __global__ __launch_bounds__(64) void k_chain_fold(const GenControl *ctl, GenControl *heads,
                                                           int n_list_faces, const rpx_face_t *faces,
                                                           const rpx_material_t *materials) {
    const GenControl &c = ctl[blockIdx.x];
    GenControl &h = heads[blockIdx.x];
    __shared__ uint32_t lds_total[64];
    uint32_t mine = 0;
    for (int k = threadIdx.x; k < n_list_faces; k += 64) {
        uint32_t t = 0;
        for (int q = 0; q < RPX_SHARDS; ++q) t += c.face_shards[q][k];
        h.face_totals[k] = t;
        const int m = faces[k].material;
        if (m >= 0 && material_emits(materials[m])) mine += t;
    }
    // every ray that hit a face is counted once, on that face: the rays whose
    // material was evaluated are the face totals of faces whose material can
    // have children (the generation kernels of a chain leave the hit shards
    // alone)
    lds_total[threadIdx.x] = mine;
    __syncthreads();
    if (threadIdx.x < RPX_SHARDS) h.hit_shards[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        uint32_t all = 0;
        for (int k = 0; k < 64; ++k) all += lds_total[k];
        h.hit_shards[0] = all;
        uint64_t parents = 0;   // gausslet generations: gausslets with children
        for (int q = 0; q < RPX_SHARDS; ++q) parents += c.parent_shards[q];
        h.parents = parents;
        h.error = c.error;
    }
}

// Pinned, device-mapped control heads, one per generation of a chain, per
// host thread; grown on demand and held until the thread ends.
struct PinnedHeads {
    GenControl *host = nullptr, *dev = nullptr;
    int n = 0;
    ~PinnedHeads() {
        if (host) (void)hipHostFree(host);
    }
};
static inline GenControl *pinned_heads(int n, GenControl **dev) {
    thread_local PinnedHeads h;
    if (h.n < n) {
        if (h.host) (void)hipHostFree(h.host);
        h.host = h.dev = nullptr;
        h.n = 0;
        GenControl *p = nullptr;
        void *d = nullptr;
        const int want = std::max(n, 64);
        if (hipHostMalloc((void **)&p, sizeof(GenControl) * want, hipHostMallocMapped | hipHostMallocPortable) !=
            hipSuccess)
            return nullptr;
        if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess) {
            (void)hipHostFree(p);
            return nullptr;
        }
        h.host = p;
        h.dev = (GenControl *)d;
        h.n = want;
    }
    *dev = h.dev;
    return h.host;
}

// What a generation launch of a chain needs besides the kernel's own
// arguments.  Tile order: with by_launch, a workgroup's tile is tile_base +
// blockIdx.x (a generation's second launch starts at the first launch's
// grid) and its look-back waits are bounded by lookback_ticks; otherwise
// tiles are handed out by the generation's atomic counter.
struct ChainGen {
    GenLink *link;
    GenControl *ctl;
    uint64_t *tile_status;
    uint64_t tag;
    GenControl *head_out;
    int32_t in_pool, store_children;
    uint32_t tile_base;
    int32_t by_launch;
    uint64_t lookback_ticks;
};

// Ticks of the device's real-time counter in `ms` milliseconds (the counter
// s_memrealtime reads; its rate is the device's wall-clock rate).
static inline uint64_t realtime_ticks(double ms) {
    thread_local int dev_cached = -1;
    thread_local uint64_t khz = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (dev != dev_cached) {
        int rate = 0;
        if (hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, dev) != hipSuccess || rate <= 0)
            rate = 100000;   // 100 MHz
        khz = (uint64_t)rate;
        dev_cached = dev;
    }
    return std::max<uint64_t>((uint64_t)(ms * (double)khz), 2);
}

struct ChainSpec {
    uint64_t n_in, pool_rays, max_children;
    uint32_t block = RPX_BLOCK;               // rays per tile of the generation kernel (<= the workspace's)
    bool bytes_per_parent = false;            // bytes_hit per gausslet with children (gausslet chains)
    int32_t max_generations;
    int timing_slot;                          // RPX_TIMING_* of the generation kernel
    double bytes_in, bytes_hit, bytes_child;  // algorithmic bytes per input ray / hit / child
};

static inline rpx_status_t check_chain(const ChainSpec &c, const rpx_scene_t *scene, void *workspace,
                                       uint64_t workspace_bytes) {
    const uint64_t max_gen = std::max<uint64_t>(c.n_in, c.pool_rays);
    if (max_gen > 0xFFFFFFFFull) return rpx_set_error(RPX_E_UNSUPPORTED, "more than 2^32-1 rays in one generation");
    if (!workspace || workspace_bytes < chain_workspace_bytes(max_gen, c.max_generations))
        return rpx_set_error(RPX_E_INVALID, "workspace too small");
    if ((int)scene->face_global_idx.size() < scene->view.n_list_faces)
        return rpx_set_error(RPX_E_INVALID, "scene face table");
    return RPX_OK;
}

// Status of one chain attempt whose launch-order look-back outlasted its
// bound (never returned to callers).
static constexpr rpx_status_t RPX_CHAIN_RETRY = -1;

// Default bound on one look-back wait with launch-order tiles, in ms.  A
// tile waits only for tiles of workgroups dispatched before it, whose work
// takes microseconds to a few milliseconds; the bound is met only if the
// hardware did not dispatch in workgroup order.
#define RPX_LOOKBACK_BOUND_MS 50.0

template <class Launch>
static rpx_status_t run_chain_once(const ChainSpec &c, const rpx_scene_t *scene, void *workspace, Launch &launch,
                                   uint64_t *counts, uint64_t *offsets, int32_t *n_traced, uint32_t *face_hits,
                                   hipStream_t stream, bool by_launch);

// Runs the chain; `launch(gen, grid, stream)` launches one generation's
// kernel (grid workgroups) with the chain fields of `gen` and returns the
// launch error.  counts / offsets / n_traced / face_hits (host) as
// rpx_trace_rays documents them.
//
// Tiles are taken in launch order (tile = workgroup id): taking them from an
// atomic counter costs one same-address device atomic per tile, and those
// serialise at about 12 ns each, a floor of 0.19 ms per 16,384-tile
// generation.  Launch order needs the hardware to start workgroups in id
// order for the look-back's waits to end; each wait is therefore bounded,
// and a chain in which one ran out is traced again from its input with
// ticket order, which guarantees progress whatever the dispatch order.  (A
// run-out wait leaves positions below the true ones, inside every buffer,
// and the rerun rewrites every generation and the input's lengths and end
// faces.)
template <class Launch>
static rpx_status_t run_chain(const ChainSpec &c, const rpx_scene_t *scene, void *workspace, Launch &&launch,
                              uint64_t *counts, uint64_t *offsets, int32_t *n_traced, uint32_t *face_hits,
                              hipStream_t stream) {
    // an attempt adds its hit totals only once it has ended without a run-out wait
    rpx_status_t st = run_chain_once(c, scene, workspace, launch, counts, offsets, n_traced, face_hits, stream, true);
    if (st != RPX_CHAIN_RETRY) return st;
    rpx_note_tile_order_fallback();
    return run_chain_once(c, scene, workspace, launch, counts, offsets, n_traced, face_hits, stream, false);
}

template <class Launch>
static rpx_status_t run_chain_once(const ChainSpec &c, const rpx_scene_t *scene, void *workspace, Launch &launch,
                                   uint64_t *counts, uint64_t *offsets, int32_t *n_traced, uint32_t *face_hits,
                                   hipStream_t stream, bool by_launch) {
    const uint64_t ticks = rpx_lookback_ticks_override() ? rpx_lookback_ticks_override()
                                                         : realtime_ticks(RPX_LOOKBACK_BOUND_MS);
    *n_traced = 0;
    const int32_t L = c.max_generations;
    const uint64_t mc = c.max_children, max_gen = std::max<uint64_t>(c.n_in, c.pool_rays);
    GenControl *heads_dev = nullptr;
    GenControl *heads = pinned_heads(L + 1, &heads_dev);
    if (!heads) return rpx_set_error(RPX_E_HIP, "pinned control heads could not be allocated");
    const ChainWs cw = carve_chain(workspace, max_gen, L);
    const bool timing = rpx_timing_is_on();
    static thread_local std::vector<hipEvent_t> evs;
    // event 2g: generation g starts, 2g+1: it has ended; 2L: the hit fold has ended
    while (evs.size() <= 2 * (size_t)L) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return rpx_set_error(RPX_E_HIP, "event creation failed");
        evs.push_back(e);
    }
    {
        const uint64_t tiles = chain_tiles(max_gen);
        const uint64_t words = std::max<uint64_t>(tiles, sizeof(GenControl) / 4 * (uint64_t)(L + 1));
        hipLaunchKernelGGL(k_chain_init, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, stream, cw, tiles, L,
                           c.n_in, heads_dev);
        RPX_HIP_TRY(hipGetLastError());
    }
    // counts[g] is exact once known[g], offsets[g] with it; grid[g]: tiles of
    // generation g's first launch; a generation is closed (second launch for
    // tiles past its grid, then its end event) once its size is known and
    // before the next generation is launched.  Heads are read in order.
    std::vector<char> known(L + 2, 0), closed(L + 2, 0);
    std::vector<uint64_t> grid(L + 2, 0);
    known[0] = 1;
    int launched = 0, synced = 0;
    rpx_status_t result = RPX_OK;
    auto tiles_of = [&](uint64_t n) { return (n + c.block - 1) / c.block; };
    auto gen = [&](int g) {
        ChainGen cg;
        cg.link = cw.link + g;
        cg.ctl = cw.ctl + g;
        cg.tile_status = cw.tile_status;
        cg.tag = (uint64_t)g + 1;
        cg.head_out = heads_dev + g;
        cg.in_pool = g > 0;
        cg.store_children = g + 1 < L ? 1 : 0;
        cg.tile_base = 0;
        cg.by_launch = by_launch ? 1 : 0;
        cg.lookback_ticks = by_launch ? ticks : 0;
        return cg;
    };
    auto close_gen = [&](int g) -> rpx_status_t {
        if (closed[g]) return RPX_OK;
        const uint64_t need = tiles_of(counts[g]);
        if (need > grid[g]) {
            ChainGen cg = gen(g);
            cg.tile_base = (uint32_t)grid[g];
            RPX_HIP_TRY(launch(cg, need - grid[g], stream));
        }
        RPX_HIP_TRY(hipEventRecord(evs[2 * g + 1], stream));
        closed[g] = 1;
        return RPX_OK;
    };
    auto read_heads_through = [&](int last) -> rpx_status_t {
        for (int g = synced; g <= last; ++g) {
            if (rpx_status_t st = close_gen(g)) return st;
            RPX_HIP_TRY(hipEventSynchronize(evs[2 * g + 1]));
            counts[g + 1] = heads[g].children;
            offsets[g + 1] = g == 0 ? 0 : offsets[g] + counts[g];
            known[g + 1] = 1;
            synced = g + 1;
        }
        return RPX_OK;
    };
    for (int g = 0; g < L; ++g) {
        // the size of generation g-1 bounds generation g: learn it from
        // generation g-2's head and close generation g-1
        if (g >= 2 && !known[g - 1])
            if (rpx_status_t st = read_heads_through(g - 2)) return st;
        if (g >= 1 && known[g - 1] && !closed[g - 1])
            if (rpx_status_t st = close_gen(g - 1)) return st;
        if (known[g] && counts[g] == 0) break;
        if (g >= 1 && counts[g - 1] == 0) break;
        uint64_t bound = known[g] ? counts[g] : mc * counts[g - 1];
        if (bound > max_gen) {   // beyond the workspace's tiles: take the exact size
            if (rpx_status_t st = read_heads_through(g - 1)) return st;
            if (counts[g] == 0) break;
            bound = counts[g];
        }
        if (g + 1 < L) {
            // where the children could end: generation g's own place (from
            // offsets[g], exact once counts[g-1] is) plus its worst case
            auto need_for = [&](uint64_t b) {
                const uint64_t start = g == 0 ? 0 : (g == 1 ? 0 : offsets[g - 1] + counts[g - 1]) + b;
                return start + mc * b;
            };
            uint64_t need = need_for(bound);
            if (need > c.pool_rays && !known[g]) {
                if (rpx_status_t st = read_heads_through(g - 1)) return st;
                if (counts[g] == 0) break;
                bound = counts[g];
                need = need_for(bound);
            }
            if (need > c.pool_rays) {
                result = rpx_set_error(RPX_E_CAPACITY, "pool of " + std::to_string(c.pool_rays) + " rays cannot "
                                       "take generation " + std::to_string(g) + "'s children (" +
                                       std::to_string(need) + " at worst): it is stored and not traced");
                break;
            }
        }
        // exact when the size is known, else the last observed
        // children-per-parent ratio applied to the parent generation
        uint64_t guess;
        if (known[g]) {
            guess = counts[g];
        } else {
            const double ratio = (g >= 2 && counts[g - 2] > 0) ? (double)counts[g - 1] / (double)counts[g - 2] : 1.0;
            guess = std::min<uint64_t>(bound, (uint64_t)std::ceil(ratio * (double)counts[g - 1]));
        }
        grid[g] = std::max<uint64_t>(tiles_of(guess), 1);
        if (timing) RPX_HIP_TRY(hipEventRecord(evs[2 * g], stream));
        RPX_HIP_TRY(launch(gen(g), grid[g], stream));
        launched = g + 1;
        if (known[g])
            if (rpx_status_t st = close_gen(g)) return st;
    }
    if (launched > 0) {
        // the size of the last launched generation (from its parent's head)
        // closes it; then every generation's hit totals, then every head
        if (launched >= 2 && synced < launched - 1)
            if (rpx_status_t st = read_heads_through(launched - 2)) return st;
        if (rpx_status_t st = close_gen(launched - 1)) return st;
        hipLaunchKernelGGL(k_chain_fold, dim3((unsigned)launched), dim3(64), 0, stream, (const GenControl *)cw.ctl,
                           heads_dev, scene->view.n_list_faces, (const rpx_face_t *)scene->view.faces,
                           (const rpx_material_t *)scene->view.materials);
        RPX_HIP_TRY(hipGetLastError());
        RPX_HIP_TRY(hipEventRecord(evs[2 * L], stream));
        if (launched > synced)
            if (rpx_status_t st = read_heads_through(launched - 1)) return st;
        RPX_HIP_TRY(hipEventSynchronize(evs[2 * L]));
        for (int g = 0; g < launched; ++g)
            if (heads[g].error) {
                if (by_launch) return RPX_CHAIN_RETRY;
                return rpx_set_error(RPX_E_HIP, "generation scan: a look-back wait exceeded its poll limit");
            }
    }
    int traced = 0;
    for (int g = 0; g < launched; ++g) {
        if (counts[g] == 0) break;
        traced = g + 1;
    }
    *n_traced = traced;
    for (int g = 0; g < traced; ++g)
        for (int k = 0; k < scene->view.n_list_faces; ++k)
            face_hits[scene->face_global_idx[k]] += heads[g].face_totals[k];
    if (timing) {
        for (int g = 0; g < traced; ++g) {
            float ms = 0;
            RPX_HIP_TRY(hipEventElapsedTime(&ms, evs[2 * g], evs[2 * g + 1]));
            uint64_t hits = 0;
            for (int s2 = 0; s2 < RPX_SHARDS; ++s2) hits += heads[g].hit_shards[s2];
            if (c.bytes_per_parent) hits = heads[g].parents;
            const bool stored = g + 1 < L;
            rpx_timing_add(c.timing_slot, ms, (double)counts[g] * c.bytes_in + (double)hits * c.bytes_hit +
                                                  (stored ? (double)counts[g + 1] * c.bytes_child : 0.0));
        }
    }
    return result;
}

static inline uint32_t scene_variant(const rpx_scene_t *s) {
    if ((s->features & ~(uint32_t)V_SPHERES) == 0) return V_SPHERES;
    if ((s->features & ~(uint32_t)V_LENS) == 0) return V_LENS;
    return V_ALL;
}

struct EventSet {
    hipEvent_t e[5];
    bool ok = false;
    EventSet() {
        ok = true;
        for (auto &x : e)
            if (hipEventCreate(&x) != hipSuccess) ok = false;
    }
    ~EventSet() {
        for (auto &x : e) (void)hipEventDestroy(x);
    }
};


}  // namespace rpx
