// Continuous-batching trace server.
//
// Every active job owns a float64 ray generation in HBM.  One server step
// traces the current generation of all active jobs together, and then up to
// three more without a host round trip (each laid out on the device from
// the previous one's child counts, k_batch_relayout): the jobs' rays are
// laid end to end as tiles of RPX_BLOCK rays, and one launch of one
// workgroup per tile takes tile numbers from an atomic counter.  A workgroup
// stages the scene of the job its tile belongs to in LDS, traces the tile as
// k_generation does (same face, orientation and material code), places the
// children with the decoupled look-back scan restricted to the job's own
// tiles, and writes the moments of the tile's detector hits; a second kernel
// sums them per job in a fixed order, so a job's results are bit-identical
// whether it runs alone or among others.  The host sequence of a step -- job
// table upload, status clear, the tile kernel, the moment sums, counts
// download -- is one HIP graph, built once per kernel variant and replayed
// every step with the tables rewritten in pinned host memory and the grids of
// its two tile-sized kernels reset (launch_step enqueues the same work node by
// node, when graph launch is turned off; the graph measured faster).
//
// Admission is first come, first served while the job's buffers (its input
// generation plus max_children times as many children) fit the HBM budget;
// the first waiting job is always admitted when nothing is active.
//
// Job buffers come from a heap of the server's own (rpx_heap.h): chunks of
// device memory carved best-fit, freed blocks merged with their neighbours,
// a chunk returned to the driver only when the chunks exceed the budget, so
// steady serving makes no allocator call.  (Allocating from a HIP memory
// pool instead was measured to take tens of milliseconds per call, and up to
// 1.8 s, once a process had created and destroyed a few pools; plain
// hipMalloc / hipFree per buffer wait for the device and cost up to
// milliseconds each: one such call stalls every job in the step.)
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <string.h>
#include <deque>
#include <string>
#include <vector>
#include "rpx_materials.h"
#include "rpx_host.h"
#include "rpx_gen.h"
#include "rpx_heap.h"

using namespace rpx;

namespace {

constexpr uint64_t RAY_BYTES = RPX_N_F64_FIELDS * 8 + RPX_N_U32_FIELDS * 4;   // 188

struct BatchJob {
    SceneView S;
    rpx_rays_t in, out;
    uint64_t n_in;
    uint64_t tile_begin, tile_end;   // global tiles [begin, end)
    double max_length;
    WlTable wl;
    int32_t detector;
    int32_t gens;                    // generations this step may trace (1..RPX_SERVE_MAX_GENS)
    int32_t max_children;
    int32_t pad;
    double *moments;                 // device [6], accumulated over generations
    uint64_t *counts;                // device [4] per generation of the step: children, hits, parents, traced (0/1)
};

struct BatchControl {
    uint64_t total_tiles;
    uint64_t n_jobs;
    uint64_t error;
    uint32_t tile_counter;
    uint32_t next_block;             // the generation of the step the next relayout lays out (1, 2, ...)
};

// Generations one server step can trace per job, and the count words each
// generation of a step has per job.
#define RPX_SERVE_MAX_GENS 4
#define RPX_SERVE_COUNT_WORDS (4 * RPX_SERVE_MAX_GENS)

__global__ __launch_bounds__(256) void k_batch_clear(uint64_t *status, uint64_t max_tiles, BatchJob *jobs,
                                                      BatchControl *ctl) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < max_tiles && k < ctl->total_tiles) status[k] = 0;
    if (k < ctl->n_jobs) {
        uint64_t *c = jobs[k].counts;
        c[0] = 0; c[1] = 0; c[2] = 0;
    }
}

// Per-job sum of the tiles' detector moments in a fixed order (each thread a
// strided run of tiles, then a fixed tree), so a job's result does not depend
// on which other jobs shared the launch or on workgroup timing.
__global__ __launch_bounds__(256) void k_batch_reduce(const BatchJob *jobs, const BatchControl *ctl,
                                                       const double *partials) {
    __shared__ double red[256][5];
    const uint64_t j = blockIdx.x;
    if (j >= ctl->n_jobs) return;
    const BatchJob &J = jobs[j];
    double s[5] = {0, 0, 0, 0, 0};
    for (uint64_t t = J.tile_begin + threadIdx.x; t < J.tile_end; t += 256)
        for (int q = 0; q < 5; ++q) s[q] += partials[t * 5 + q];
    for (int q = 0; q < 5; ++q) red[threadIdx.x][q] = s[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < (unsigned)w)
            for (int q = 0; q < 5; ++q) red[threadIdx.x][q] += red[threadIdx.x + w][q];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 5; ++q) J.moments[q] += red[0][q];
}

// Lays out a step's next generation on the device, from the previous one's
// child counts: a job that may trace it (gens > the generation's index in
// the step) and whose children fit, max_children each, in the buffer the
// previous generation came from traces them from the previous generation's
// output into that buffer, which is no longer needed (the two buffers of a
// job alternate); the other jobs get no tiles.  Jobs keep their order, so
// the tiles stay laid end to end.  Each job's counts move to the next block
// of four words, whose last word records whether the generation was traced.
// One workgroup: each thread takes a run of jobs.
__global__ __launch_bounds__(256) void k_batch_relayout(BatchJob *jobs, BatchControl *ctl) {
    __shared__ uint64_t run_tiles[256];
    const uint32_t n = (uint32_t)ctl->n_jobs;
    const uint32_t blk = ctl->next_block;
    const uint32_t per = (n + 255) / 256;
    const uint32_t b = threadIdx.x * per, e = min(n, b + per);
    auto traces = [blk](const BatchJob &J, uint64_t kids) {
        return J.gens > (int32_t)blk && kids > 0 && kids * (uint64_t)J.max_children <= J.in.capacity;
    };
    uint64_t mine = 0;
    for (uint32_t j = b; j < e; ++j) {
        const uint64_t kids = jobs[j].counts[0];
        if (traces(jobs[j], kids)) mine += (kids + RPX_BLOCK - 1) / RPX_BLOCK;
    }
    run_tiles[threadIdx.x] = mine;
    __syncthreads();
    uint64_t base = 0;
    for (uint32_t t = 0; t < threadIdx.x; ++t) base += run_tiles[t];
    for (uint32_t j = b; j < e; ++j) {
        BatchJob &J = jobs[j];
        const uint64_t kids = J.counts[0];
        const bool go = traces(J, kids);
        const rpx_rays_t first_in = J.in;
        J.in = J.out;
        J.out = first_in;
        J.n_in = go ? kids : 0;
        J.tile_begin = base;
        base += go ? (kids + RPX_BLOCK - 1) / RPX_BLOCK : 0;
        J.tile_end = base;
        J.counts += 4;
        J.counts[3] = go ? 1 : 0;
    }
    if (threadIdx.x == 255) {
        ctl->total_tiles = base;
        ctl->tile_counter = 0;
        ctl->next_block = blk + 1;
    }
}

template <uint32_t F>
__global__ __launch_bounds__(RPX_BLOCK) __attribute__((amdgpu_waves_per_eu(F == V_ALL ? 1 : 3)))
void k_batch(const BatchJob *jobs, BatchControl *ctl, uint64_t *status,
                                                     double *partials) {
    __shared__ SceneLds lds_scene;
    __shared__ uint32_t lds_wave[RPX_BLOCK / 64];
    __shared__ uint64_t lds_tile, lds_excl;
    __shared__ int32_t lds_job;
    __shared__ uint32_t lds_parents, lds_hits;
    __shared__ double lds_mom[RPX_BLOCK / 64][5];
    const uint64_t total = ctl->total_tiles;
    const int n_jobs = (int)ctl->n_jobs;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    {
        if (threadIdx.x == 0) {
            const uint64_t t = atomicAdd(&ctl->tile_counter, 1u);
            int j = -1;
            if (t < total) {
                int lo = 0, hi = n_jobs - 1;   // last job with tile_begin <= t
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (jobs[mid].tile_begin <= t) lo = mid;
                    else hi = mid - 1;
                }
                j = lo;
            }
            lds_tile = t;
            lds_job = j;
            lds_parents = 0;
            lds_hits = 0;
        }
        __syncthreads();
        const uint64_t tile = lds_tile;
        if (tile >= total) return;
        const int j = lds_job;
        const BatchJob &J = jobs[j];
        const SceneView S = stage_scene(J.S, lds_scene);   // ends with a barrier
        const uint64_t i = (tile - J.tile_begin) * RPX_BLOCK + threadIdx.x;
        PairEmit ce;
        bool hit = false;
        double mx = 0, my = 0, mz = 0, mr = 0, mn = 0;
        if (i < J.n_in) {
            const double *f = J.in.f64;
            const uint64_t c = J.in.capacity;
            const V3 origin = v3(f[RPX_OX * c + i], f[RPX_OY * c + i], f[RPX_OZ * c + i]);
            const V3 dir = v3(f[RPX_DX * c + i], f[RPX_DY * c + i], f[RPX_DZ * c + i]);
            double length = J.max_length;
            uint32_t end_face = RPX_NO_FACE;
            int hit_rec = -1, hit_list = -1;
            const V3 point = add(origin, mul(dir, J.max_length));
            // trace_segment_c / FaceList.intersect_c (ctracer.pyx:2093-2104, 1882-1904)
            for (int l = 0; l < S.n_face_lists; ++l) {
                const rpx_face_list_t &fl = S.face_lists[l];
                const V3 p1 = transform(fl.inv_trans, origin);
                const V3 p2 = transform(fl.inv_trans, point);
                const double slen = sep(p1, p2), zrate = (p2.z - p1.z) / slen;
                for (int k = fl.first_face; k < fl.first_face + fl.n_faces; ++k) {
                    const rpx_face_t &fc = S.faces[k];
                    if (face_out_of_reach(fc, p1.z, zrate * length, slen)) continue;
                    const double dist = face_intersect<F>(S, fc, p1, p2, 1);
                    if (fc.tolerance < dist && dist < length) {
                        length = dist;
                        end_face = (uint32_t)fc.global_idx;
                        hit_rec = k;
                        hit_list = l;
                    }
                }
            }
            J.in.f64[RPX_LENGTH * c + i] = length;
            J.in.u32[RPX_END_FACE_IDX * c + i] = end_face;
            if (hit_rec >= 0) {
                hit = true;
                const rpx_face_t &fc = S.faces[hit_rec];
                const V3 hp = add(origin, mul(dir, length));
                if ((int32_t)end_face == J.detector) {
                    mn = 1.0;
                    mx = hp.x; my = hp.y; mz = hp.z;
                    mr = hp.x * hp.x + hp.y * hp.y + hp.z * hp.z;
                }
                if (fc.material >= 0 && material_emits(S.materials[fc.material])) {
                    const rpx_face_list_t &fl = S.face_lists[hit_list];
                    Ray r = load_ray(J.in, i);
                    r.length = length;
                    r.end_face_idx = end_face;
                    const Orientation o = compute_orientation<F>(S, fl, fc, hp);
                    eval_children<F>(S.materials[fc.material], r, (uint32_t)i, hp, o, J.wl, ce);
                }
            }
        }
        const uint32_t count = (uint32_t)min(ce.n, 2);
        uint32_t block_total;
        const uint32_t local = block_scan_small(count, lds_wave, &block_total);
        {
            const uint64_t mp = __ballot(count != 0), mh = __ballot(hit);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                mn += __shfl_xor(mn, o);
                mx += __shfl_xor(mx, o);
                my += __shfl_xor(my, o);
                mz += __shfl_xor(mz, o);
                mr += __shfl_xor(mr, o);
            }
            if (lane == 0) {
                atomicAdd(&lds_parents, (uint32_t)__popcll(mp));
                atomicAdd(&lds_hits, (uint32_t)__popcll(mh));
                lds_mom[wave][0] = mn;
                lds_mom[wave][1] = mx;
                lds_mom[wave][2] = my;
                lds_mom[wave][3] = mz;
                lds_mom[wave][4] = mr;
            }
        }
        if (threadIdx.x < 64) {
            const uint64_t excl = lookback(status, tile, J.tile_begin, block_total, &ctl->error);
            if (threadIdx.x == 0) lds_excl = excl;
        }
        __syncthreads();
        const uint64_t pos = lds_excl + local;
        if (count > 0 && pos < J.out.capacity) store_ray(J.out, pos, ce.first());
        if (count > 1 && pos + 1 < J.out.capacity) store_ray(J.out, pos + 1, ce.r0);
        if (threadIdx.x == 0) {
            if (lds_hits) atomicAdd((unsigned long long *)&J.counts[1], (unsigned long long)lds_hits);
            if (lds_parents) atomicAdd((unsigned long long *)&J.counts[2], (unsigned long long)lds_parents);
            if (tile == J.tile_end - 1) J.counts[0] = lds_excl + block_total;
            for (int q = 0; q < 5; ++q) {
                double v = 0;
                for (int w = 0; w < RPX_BLOCK / 64; ++w) v += lds_mom[w][q];
                partials[tile * 5 + q] = v;
            }
        }
    }
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Buffer {
    void *p = nullptr;
    uint64_t cap = 0;
    rpx_rays_t rays() const {
        rpx_rays_t r;
        r.f64 = (double *)p;
        r.u32 = (uint32_t *)((char *)p + (uint64_t)RPX_N_F64_FIELDS * 8 * cap);
        r.para = nullptr;
        r.capacity = cap;
        return r;
    }
    uint64_t bytes() const { return cap * RAY_BYTES; }
};

struct Job {
    uint64_t id = 0;
    rpx_job_desc_t desc{};
    uint32_t variant = V_ALL;
    int max_children = 2;
    Buffer cur, next;
    uint64_t n = 0;
    int slot = -1;
    int gen = 0;
    uint64_t interactions = 0, children = 0;
    double t_submit = 0, t_admit = 0;
    int status = RPX_OK;
    bool held_for_budget = false;   // once held back at admission by the HBM budget
};

}  // namespace

// A step's graph: per generation a clear and a tile kernel whose grids are
// set before each replay (and a relayout before every generation but the
// first).
struct StepGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t exec = nullptr;
    int gens = 0;                                        // generations the graph traces
    hipGraphNode_t clear[RPX_SERVE_MAX_GENS] = {}, batch[RPX_SERVE_MAX_GENS] = {};
    void *batch_func = nullptr;
};

struct rpx_server {
    hipStream_t stream = nullptr;
    rpx::BlockHeap<DeviceChunks> heap;   // job buffers (chunks returned only above the budget, between steps)
    uint64_t budget = 0, used = 0;
    uint64_t peak_used = 0, jobs_held_for_budget = 0;
    uint32_t max_jobs = 0;
    uint64_t max_tiles = 0;
    BatchJob *d_jobs = nullptr, *h_jobs = nullptr;
    BatchControl *d_ctl = nullptr, *h_ctl = nullptr, *h_ctl_back = nullptr;
    uint64_t *d_status = nullptr;
    double *d_partials = nullptr;
    uint64_t *d_counts = nullptr, *h_counts = nullptr;
    double *d_moments = nullptr, *h_moments = nullptr;
    StepGraph graphs[RPX_SERVE_MAX_GENS][3];   // by generations per step - 1, then kernel variant
    uint32_t gens_per_step = 2;
    bool graph_launch = true;    // replay the step graph (else launch its nodes one by one)
    void *clear_args[4] = {&d_status, &max_tiles, &d_jobs, &d_ctl};
    void *batch_args[4] = {&d_jobs, &d_ctl, &d_status, &d_partials};
    void *reduce_args[3] = {&d_jobs, &d_ctl, &d_partials};
    void *relayout_args[2] = {&d_jobs, &d_ctl};
    std::deque<Job *> waiting;
    std::vector<Job *> active;
    std::vector<int> free_slots;
    std::deque<rpx_job_result_t> done;
    uint64_t next_id = 1;
    int64_t allocs_until_failure = -1;   // test hook: the allocation after this many fails; < 0 never
};

static rpx_status_t alloc_buffer(rpx_server_t *s, Buffer &b, uint64_t cap) {
    if (s->allocs_until_failure == 0) {
        s->allocs_until_failure = -1;
        b = Buffer();
        return rpx_set_error(RPX_E_HIP, "job buffer allocation failed (injected)");
    }
    if (s->allocs_until_failure > 0) --s->allocs_until_failure;
    b.cap = std::max<uint64_t>(cap, 1);
    b.p = s->heap.alloc(b.bytes());
    if (!b.p) {
        b = Buffer();
        return rpx_set_error(RPX_E_HIP, "job buffer allocation: out of device memory");
    }
    return RPX_OK;
}

// Every kernel that used the buffer has completed: the server synchronises
// its stream after each step and after an input copy.
static void free_buffer(rpx_server_t *s, Buffer &b) {
    if (b.p) (void)s->heap.free(b.p);
    b.p = nullptr;
    b.cap = 0;
}

#define RPX_GRAPH_TRY(expr)                                                                   \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) return rpx_set_error(RPX_E_HIP, std::string("step graph: ") + #expr + ": " + \
                                                                   hipGetErrorString(_e));    \
    } while (0)

template <uint32_t F>
static rpx_status_t build_graph(rpx_server_t *s, StepGraph &G, int gens) {
    RPX_GRAPH_TRY(hipGraphCreate(&G.g, 0));
    hipGraphNode_t last, node;
    RPX_GRAPH_TRY(hipGraphAddMemcpyNode1D(&node, G.g, nullptr, 0, s->d_jobs, s->h_jobs,
                                          sizeof(BatchJob) * s->max_jobs, hipMemcpyHostToDevice));
    last = node;
    RPX_GRAPH_TRY(hipGraphAddMemcpyNode1D(&node, G.g, &last, 1, s->d_ctl, s->h_ctl, sizeof(BatchControl),
                                          hipMemcpyHostToDevice));
    last = node;
    auto add_kernel = [&](void *func, dim3 grid, dim3 block, void **args, hipGraphNode_t *out) -> hipError_t {
        hipKernelNodeParams p{};
        p.func = func;
        p.gridDim = grid;
        p.blockDim = block;
        p.kernelParams = args;
        hipError_t e = hipGraphAddKernelNode(&node, G.g, &last, 1, &p);
        if (e == hipSuccess) {
            last = node;
            if (out) *out = node;
        }
        return e;
    };
    // per generation: (after the first, the device relayout) status and
    // counts clear, the tile kernel, the moment sums
    for (int k = 0; k < gens; ++k) {
        if (k > 0)
            RPX_GRAPH_TRY(add_kernel((void *)k_batch_relayout, dim3(1), dim3(256), s->relayout_args, nullptr));
        RPX_GRAPH_TRY(add_kernel((void *)k_batch_clear, dim3(1), dim3(256), s->clear_args, &G.clear[k]));
        RPX_GRAPH_TRY(add_kernel((void *)k_batch<F>, dim3(1), dim3(RPX_BLOCK), s->batch_args, &G.batch[k]));
        RPX_GRAPH_TRY(add_kernel((void *)k_batch_reduce, dim3(s->max_jobs), dim3(256), s->reduce_args, nullptr));
    }
    G.gens = gens;
    RPX_GRAPH_TRY(hipGraphAddMemcpyNode1D(&node, G.g, &last, 1, s->h_counts, s->d_counts,
                                          sizeof(uint64_t) * RPX_SERVE_COUNT_WORDS * s->max_jobs,
                                          hipMemcpyDeviceToHost));
    last = node;
    RPX_GRAPH_TRY(hipGraphAddMemcpyNode1D(&node, G.g, &last, 1, s->h_moments, s->d_moments,
                                          sizeof(double) * 6 * s->max_jobs, hipMemcpyDeviceToHost));
    last = node;
    RPX_GRAPH_TRY(hipGraphAddMemcpyNode1D(&node, G.g, &last, 1, s->h_ctl_back, s->d_ctl, sizeof(BatchControl),
                                          hipMemcpyDeviceToHost));
    RPX_GRAPH_TRY(hipGraphInstantiate(&G.exec, G.g, nullptr, nullptr, 0));
    G.batch_func = (void *)k_batch<F>;
    return RPX_OK;
}

// Point the tile-sized kernels of the graph at this step's tile counts: the
// first generation's exact count, and bounds on the later ones' (their
// workgroups past the relayout's count leave at once).
static rpx_status_t size_graph(rpx_server_t *s, StepGraph &G, const uint64_t *tiles, uint64_t n_jobs) {
    hipKernelNodeParams p{};
    auto set = [&](hipGraphNode_t clear, hipGraphNode_t batch, uint64_t t) -> rpx_status_t {
        p.func = (void *)k_batch_clear;
        p.gridDim = dim3((unsigned)((std::max(t, n_jobs) + 255) / 256));
        p.blockDim = dim3(256);
        p.kernelParams = s->clear_args;
        RPX_GRAPH_TRY(hipGraphExecKernelNodeSetParams(G.exec, clear, &p));
        p.func = G.batch_func;
        p.gridDim = dim3((unsigned)std::max<uint64_t>(t, 1));
        p.blockDim = dim3(RPX_BLOCK);
        p.kernelParams = s->batch_args;
        RPX_GRAPH_TRY(hipGraphExecKernelNodeSetParams(G.exec, batch, &p));
        return RPX_OK;
    };
    for (int k = 0; k < G.gens; ++k)
        if (rpx_status_t st = set(G.clear[k], G.batch[k], tiles[k])) return st;
    return RPX_OK;
}

// The step graph's work enqueued node by node on the server's stream: the
// job table (only the batched jobs' entries) and control block up, per
// generation the relayout (after the first), clear, tile kernel and moment
// sums, then the counts, moments and control block back.
template <uint32_t F>
static rpx_status_t launch_step(rpx_server_t *s, int gens, const uint64_t *tiles, uint64_t n_jobs) {
    hipStream_t st = s->stream;
    RPX_HIP_TRY(hipMemcpyAsync(s->d_jobs, s->h_jobs, sizeof(BatchJob) * n_jobs, hipMemcpyHostToDevice, st));
    RPX_HIP_TRY(hipMemcpyAsync(s->d_ctl, s->h_ctl, sizeof(BatchControl), hipMemcpyHostToDevice, st));
    for (int k = 0; k < gens; ++k) {
        if (k > 0) hipLaunchKernelGGL(k_batch_relayout, dim3(1), dim3(256), 0, st, s->d_jobs, s->d_ctl);
        const uint64_t t = tiles[k];
        hipLaunchKernelGGL(k_batch_clear, dim3((unsigned)((std::max(t, n_jobs) + 255) / 256)), dim3(256), 0, st,
                           s->d_status, s->max_tiles, s->d_jobs, s->d_ctl);
        hipLaunchKernelGGL((k_batch<F>), dim3((unsigned)std::max<uint64_t>(t, 1)), dim3(RPX_BLOCK), 0, st,
                           s->d_jobs, s->d_ctl, s->d_status, s->d_partials);
        hipLaunchKernelGGL(k_batch_reduce, dim3((unsigned)n_jobs), dim3(256), 0, st, s->d_jobs, s->d_ctl,
                           s->d_partials);
    }
    RPX_HIP_TRY(hipGetLastError());
    RPX_HIP_TRY(hipMemcpyAsync(s->h_counts, s->d_counts, sizeof(uint64_t) * RPX_SERVE_COUNT_WORDS * s->max_jobs,
                               hipMemcpyDeviceToHost, st));
    RPX_HIP_TRY(hipMemcpyAsync(s->h_moments, s->d_moments, sizeof(double) * 6 * s->max_jobs, hipMemcpyDeviceToHost,
                               st));
    RPX_HIP_TRY(hipMemcpyAsync(s->h_ctl_back, s->d_ctl, sizeof(BatchControl), hipMemcpyDeviceToHost, st));
    return RPX_OK;
}

extern "C" double rpx_server_clock(void) { return now_s(); }

extern "C" rpx_status_t rpx_server_create(uint64_t hbm_budget_bytes, uint32_t max_jobs, uint64_t max_rays_in_flight,
                                          void *stream, rpx_server_t **out) {
    (void)stream;   // the server runs its graph on a stream of its own
    if (!out) return rpx_set_error(RPX_E_INVALID, "out is null");
    *out = nullptr;
    if (max_jobs == 0 || max_jobs > 4096) return rpx_set_error(RPX_E_INVALID, "max_jobs must be 1..4096");
    if (max_rays_in_flight == 0) return rpx_set_error(RPX_E_INVALID, "max_rays_in_flight is 0");
    rpx_server_t *s = new rpx_server_t();
    s->budget = hbm_budget_bytes;
    s->heap.set_limits(hbm_budget_bytes, std::min<uint64_t>(hbm_budget_bytes, 256ull << 20));
    s->max_jobs = max_jobs;
    s->max_tiles = (max_rays_in_flight + RPX_BLOCK - 1) / RPX_BLOCK + max_jobs;
    auto fail = [&](rpx_status_t st) {
        rpx_server_destroy(s);
        return st;
    };
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(rpx_set_error(RPX_E_HIP, hipGetErrorString(e)));
    if (hipMalloc(&s->d_jobs, sizeof(BatchJob) * max_jobs) != hipSuccess ||
        hipMalloc(&s->d_ctl, sizeof(BatchControl)) != hipSuccess ||
        hipMalloc(&s->d_status, sizeof(uint64_t) * s->max_tiles) != hipSuccess ||
        hipMalloc(&s->d_partials, sizeof(double) * 5 * s->max_tiles) != hipSuccess ||
        hipMalloc(&s->d_counts, sizeof(uint64_t) * RPX_SERVE_COUNT_WORDS * max_jobs) != hipSuccess ||
        hipMalloc(&s->d_moments, sizeof(double) * 6 * max_jobs) != hipSuccess ||
        hipHostMalloc(&s->h_jobs, sizeof(BatchJob) * max_jobs, 0) != hipSuccess ||
        hipHostMalloc(&s->h_ctl, sizeof(BatchControl), 0) != hipSuccess ||
        hipHostMalloc(&s->h_ctl_back, sizeof(BatchControl), 0) != hipSuccess ||
        hipHostMalloc(&s->h_counts, sizeof(uint64_t) * RPX_SERVE_COUNT_WORDS * max_jobs, 0) != hipSuccess ||
        hipHostMalloc(&s->h_moments, sizeof(double) * 6 * max_jobs, 0) != hipSuccess)
        return fail(rpx_set_error(RPX_E_HIP, "server allocation failed"));
    memset(s->h_jobs, 0, sizeof(BatchJob) * max_jobs);
    memset(s->h_ctl, 0, sizeof(BatchControl));
    for (int k = (int)max_jobs - 1; k >= 0; --k) s->free_slots.push_back(k);
    *out = s;
    return RPX_OK;
}

extern "C" rpx_status_t rpx_server_destroy(rpx_server_t *s) {
    if (!s) return RPX_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (Job *j : s->active) {
        free_buffer(s, j->cur);
        free_buffer(s, j->next);
        delete j;
    }
    for (Job *j : s->waiting) {
        free_buffer(s, j->cur);
        delete j;
    }
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto &row : s->graphs)
        for (StepGraph &G : row) {
            if (G.exec) (void)hipGraphExecDestroy(G.exec);
            if (G.g) (void)hipGraphDestroy(G.g);
        }
    (void)hipFree(s->d_jobs);
    (void)hipFree(s->d_ctl);
    (void)hipFree(s->d_status);
    (void)hipFree(s->d_partials);
    (void)hipFree(s->d_counts);
    (void)hipFree(s->d_moments);
    (void)hipHostFree(s->h_jobs);
    (void)hipHostFree(s->h_ctl);
    (void)hipHostFree(s->h_ctl_back);
    (void)hipHostFree(s->h_counts);
    (void)hipHostFree(s->h_moments);
    s->heap.release_all();   // every job buffer is back and the stream synchronised
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return RPX_OK;
}

extern "C" rpx_status_t rpx_server_submit(rpx_server_t *s, const rpx_job_desc_t *d, uint64_t *job_id) {
    if (!s || !d || !job_id) return rpx_set_error(RPX_E_INVALID, "null argument");
    if (!d->scene) return rpx_set_error(RPX_E_INVALID, "job scene is null");
    if (rpx_status_t st = rpx_check_trace_scene(d->scene)) return st;
    if (!d->rays || d->rays->capacity < d->n_rays || (d->n_rays && (!d->rays->f64 || !d->rays->u32)))
        return rpx_set_error(RPX_E_INVALID, "job rays");
    if (d->n_rays > 0xFFFFFFFFull) return rpx_set_error(RPX_E_UNSUPPORTED, "more than 2^32-1 rays in a job");
    if ((d->n_rays + RPX_BLOCK - 1) / RPX_BLOCK + 1 > s->max_tiles)
        return rpx_set_error(RPX_E_CAPACITY, "job larger than the server's rays in flight");
    if (d->n_wavelengths && !d->wavelengths) return rpx_set_error(RPX_E_INVALID, "wavelength table is null");
    if (d->scene->view.max_children > 2)
        return rpx_set_error(RPX_E_UNSUPPORTED, "the server traces at most two children per ray (a multi-order "
                             "grating is traced by rpx_trace_rays)");
    // a waiting job's input copy counts against the budget from submission
    const uint64_t in_bytes = std::max<uint64_t>(d->n_rays, 1) * RAY_BYTES;
    if (s->used + in_bytes > s->budget)
        return rpx_set_error(RPX_E_CAPACITY, "the job's input would exceed the server's HBM budget");
    Job *j = new Job();
    j->id = s->next_id++;
    j->desc = *d;
    j->desc.rays = nullptr;
    j->variant = scene_variant(d->scene);
    j->max_children = std::max(1, (int)d->scene->view.max_children);
    j->n = d->n_rays;
    j->t_submit = now_s();
    rpx_status_t st = alloc_buffer(s, j->cur, j->n);
    if (st) {
        delete j;
        return st;
    }
    // field rows of the caller's SoA (its own capacity) into ours
    const rpx_rays_t src = *d->rays, dst = j->cur.rays();
    if (j->n) {
        hipError_t e = hipMemcpy2DAsync(dst.f64, dst.capacity * 8, src.f64, src.capacity * 8, j->n * 8,
                                        RPX_N_F64_FIELDS, hipMemcpyDeviceToDevice, s->stream);
        if (e == hipSuccess)
            e = hipMemcpy2DAsync(dst.u32, dst.capacity * 4, src.u32, src.capacity * 4, j->n * 4, RPX_N_U32_FIELDS,
                                 hipMemcpyDeviceToDevice, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) {
            free_buffer(s, j->cur);
            delete j;
            return rpx_set_error(RPX_E_HIP, std::string("job input copy: ") + hipGetErrorString(e));
        }
    }
    s->used += j->cur.bytes();
    s->peak_used = std::max(s->peak_used, s->used);
    s->waiting.push_back(j);
    *job_id = j->id;
    return RPX_OK;
}

static void finish(rpx_server_t *s, Job *j, double t) {
    rpx_job_result_t r{};
    r.job_id = j->id;
    r.status = j->status;
    r.generations = j->gen;
    r.interactions = j->interactions;
    r.children = j->children;
    if (j->slot >= 0 && j->gen > 0)   // moments come back with every launch the job was in
        for (int q = 0; q < 5; ++q) r.moments[q] = s->h_moments[j->slot * 6 + q];
    r.t_submit = j->t_submit;
    r.t_admit = j->t_admit;
    r.t_finish = t;
    s->done.push_back(r);
    s->used -= std::min(s->used, j->cur.bytes() + j->next.bytes());
    free_buffer(s, j->cur);
    free_buffer(s, j->next);
    if (j->slot >= 0) s->free_slots.push_back(j->slot);
    delete j;
}

extern "C" rpx_status_t rpx_server_step(rpx_server_t *s, rpx_job_result_t *finished, uint32_t max_finished,
                                        uint32_t *n_finished) {
    if (!s || !n_finished) return rpx_set_error(RPX_E_INVALID, "null argument");
    *n_finished = 0;
    // admission, first come first served
    uint64_t tiles_active = 0;
    for (Job *j : s->active) tiles_active += (j->n + RPX_BLOCK - 1) / RPX_BLOCK;
    while (!s->waiting.empty() && !s->free_slots.empty()) {
        Job *j = s->waiting.front();
        // its input is already counted; the children buffer is what it adds
        const uint64_t need = std::max<uint64_t>((uint64_t)j->max_children * j->n, 1) * RAY_BYTES;
        const uint64_t tiles = (j->n + RPX_BLOCK - 1) / RPX_BLOCK;
        if (!s->active.empty() && (s->used + need > s->budget || tiles_active + tiles > s->max_tiles)) {
            if (s->used + need > s->budget && !j->held_for_budget) {
                j->held_for_budget = true;
                ++s->jobs_held_for_budget;
            }
            break;
        }
        s->waiting.pop_front();
        j->slot = s->free_slots.back();
        s->free_slots.pop_back();
        j->t_admit = now_s();
        rpx_status_t st = alloc_buffer(s, j->next, (uint64_t)j->max_children * j->n);
        if (st == RPX_OK) {
            hipError_t e = hipMemsetAsync(s->d_moments + j->slot * 6, 0, sizeof(double) * 6, s->stream);
            if (e != hipSuccess) st = rpx_set_error(RPX_E_HIP, std::string("moments clear: ") + hipGetErrorString(e));
        }
        if (st != RPX_OK) {            // the job ends here with the error; its slot and buffers go back
            j->status = st;
            finish(s, j, now_s());
            continue;
        }
        s->used += j->next.bytes();
        s->peak_used = std::max(s->peak_used, s->used);
        tiles_active += tiles;
        s->active.push_back(j);
    }
    // jobs with no rays or no generations left finish without a launch
    for (size_t k = 0; k < s->active.size();) {
        Job *j = s->active[k];
        if (j->n == 0 || j->gen >= j->desc.recursion_limit) {
            s->active.erase(s->active.begin() + k);
            finish(s, j, now_s());
        } else {
            ++k;
        }
    }
    // a job whose generation alone exceeds the status table cannot run
    for (size_t k = 0; k < s->active.size();) {
        Job *j = s->active[k];
        if ((j->n + RPX_BLOCK - 1) / RPX_BLOCK > s->max_tiles) {
            j->status = RPX_E_CAPACITY;
            s->active.erase(s->active.begin() + k);
            finish(s, j, now_s());
        } else {
            ++k;
        }
    }
    // this step's launch: active jobs in admission order while their tiles
    // fit the status table; the others wait for the next step
    std::vector<Job *> batch;
    if (!s->active.empty()) {
        uint64_t tile = 0;
        uint32_t variant = V_SPHERES;
        for (size_t k = 0; k < s->active.size(); ++k) {
            Job *j = s->active[k];
            const uint64_t tiles = (j->n + RPX_BLOCK - 1) / RPX_BLOCK;
            if (tile + tiles > s->max_tiles) continue;
            BatchJob &b = s->h_jobs[batch.size()];
            batch.push_back(j);
            b.S = j->desc.scene->view;
            b.in = j->cur.rays();
            b.out = j->next.rays();
            b.n_in = j->n;
            b.tile_begin = tile;
            tile += tiles;
            b.tile_end = tile;
            b.max_length = (double)j->desc.max_length;
            b.wl.wl = j->desc.wavelengths;
            b.wl.n = j->desc.wavelengths ? j->desc.n_wavelengths : 0;
            b.detector = j->desc.detector_face;
            b.max_children = j->max_children;
            // generations this step may trace: the job's remaining ones, up
            // to the server's setting; the device decides whether each later
            // generation's children fit
            b.gens = (int32_t)std::min<int64_t>(s->gens_per_step, (int64_t)j->desc.recursion_limit - j->gen);
            b.moments = s->d_moments + j->slot * 6;
            b.counts = s->d_counts + j->slot * RPX_SERVE_COUNT_WORDS;
            if (j->variant == V_ALL || (j->variant == V_LENS && variant == V_SPHERES)) variant = j->variant;
        }
        // tiles of each generation of the step: the first's exact, the later
        // ones' bounded by the buffer they are written into (the job's two
        // buffers alternate, the second generation writing into `cur`), over
        // max_children; the step traces as many generations as any job may
        // and the status table holds
        uint64_t tiles[RPX_SERVE_MAX_GENS] = {tile};
        int gens = 1;
        for (int g = 1; g < (int)s->gens_per_step; ++g) {
            uint64_t t = 0;
            for (size_t k = 0; k < batch.size(); ++k) {
                if (s->h_jobs[k].gens <= g) continue;
                const Buffer &into = (g % 2 == 1) ? batch[k]->cur : batch[k]->next;
                t += (into.cap / std::max(1, batch[k]->max_children) + RPX_BLOCK - 1) / RPX_BLOCK;
            }
            if (t == 0 || t > s->max_tiles) break;
            tiles[g] = t;
            gens = g + 1;
        }
        for (size_t k = 0; k < batch.size(); ++k) s->h_jobs[k].gens = std::min(s->h_jobs[k].gens, (int32_t)gens);
        s->h_ctl->total_tiles = tile;
        s->h_ctl->n_jobs = batch.size();
        s->h_ctl->error = 0;
        s->h_ctl->tile_counter = 0;
        s->h_ctl->next_block = 1;
        const int vi = variant == V_SPHERES ? 0 : variant == V_LENS ? 1 : 2;
        if (s->graph_launch) {
            StepGraph &G = s->graphs[gens - 1][vi];
            if (!G.exec) {
                rpx_status_t st = vi == 0 ? build_graph<V_SPHERES>(s, G, gens)
                                : vi == 1 ? build_graph<V_LENS>(s, G, gens)
                                          : build_graph<V_ALL>(s, G, gens);
                if (st) return st;
            }
            rpx_status_t st = size_graph(s, G, tiles, batch.size());
            if (st) return st;
            RPX_HIP_TRY(hipGraphLaunch(G.exec, s->stream));
        } else {
            rpx_status_t st = vi == 0 ? launch_step<V_SPHERES>(s, gens, tiles, batch.size())
                            : vi == 1 ? launch_step<V_LENS>(s, gens, tiles, batch.size())
                                      : launch_step<V_ALL>(s, gens, tiles, batch.size());
            if (st) return st;
        }
        RPX_HIP_TRY(hipStreamSynchronize(s->stream));
        if (s->h_ctl_back->error) {
            // the launch's counts and moments cannot be trusted: every job in
            // it ends with the error rather than being retried on moments
            // that already hold this step's partial sums
            const double t = now_s();
            std::vector<Job *> still;
            for (Job *j : s->active) {
                if (std::find(batch.begin(), batch.end(), j) == batch.end()) {
                    still.push_back(j);
                } else {
                    j->status = RPX_E_HIP;
                    finish(s, j, t);
                }
            }
            s->active.swap(still);
            return rpx_set_error(RPX_E_HIP, "batched generation: a look-back wait exceeded its poll limit; the "
                                            "jobs of that launch finished with RPX_E_HIP");
        }
        const double t = now_s();
        std::vector<Job *> still;
        size_t b = 0;
        for (Job *j : s->active) {
            if (b >= batch.size() || batch[b] != j) {   // not in this launch
                still.push_back(j);
                continue;
            }
            ++b;
            // the first generation went from `cur` into `next`; each later one
            // traced went back the other way
            const uint64_t *c = s->h_counts + j->slot * RPX_SERVE_COUNT_WORDS;
            j->interactions += c[1];
            j->children += c[0];
            j->gen += 1;
            j->n = c[0];
            bool in_cur = false;
            for (int g = 1; g < gens && c[4 * g + 3] == 1; ++g) {
                j->interactions += c[4 * g + 1];
                j->children += c[4 * g];
                j->gen += 1;
                j->n = c[4 * g];
                in_cur = !in_cur;
            }
            // `cur` holds the job's current generation, `next` a free buffer
            if (!in_cur) std::swap(j->cur, j->next);
            if (j->n > j->cur.cap) {
                j->status = RPX_E_CAPACITY;
                j->n = 0;
            }
            if (j->n == 0 || j->gen >= j->desc.recursion_limit) {
                finish(s, j, t);
                continue;
            }
            const uint64_t need = (uint64_t)j->max_children * j->n;
            if (j->next.p && j->next.cap < need) {
                s->used -= std::min(s->used, j->next.bytes());
                free_buffer(s, j->next);
            }
            if (!j->next.p) {
                rpx_status_t st = alloc_buffer(s, j->next, need);
                if (st != RPX_OK) {            // the job ends with the error; the others carry on
                    j->status = st;
                    finish(s, j, t);
                    continue;
                }
                s->used += j->next.bytes();
                s->peak_used = std::max(s->peak_used, s->used);
            }
            still.push_back(j);
        }
        s->active.swap(still);
    }
    while (!s->done.empty() && *n_finished < max_finished && finished) {
        finished[(*n_finished)++] = s->done.front();
        s->done.pop_front();
    }
    return RPX_OK;
}

extern "C" rpx_status_t rpx_server_set_generations_per_step(rpx_server_t *s, uint32_t generations) {
    if (!s) return rpx_set_error(RPX_E_INVALID, "server is null");
    if (generations < 1 || generations > RPX_SERVE_MAX_GENS)
        return rpx_set_error(RPX_E_INVALID, "generations per step must be 1.." + std::to_string(RPX_SERVE_MAX_GENS));
    s->gens_per_step = generations;
    return RPX_OK;
}

extern "C" rpx_status_t rpx_server_fail_allocations_after(rpx_server_t *s, int64_t n_allocations) {
    if (!s) return rpx_set_error(RPX_E_INVALID, "server is null");
    s->allocs_until_failure = n_allocations;
    return RPX_OK;
}

extern "C" rpx_status_t rpx_server_stats(const rpx_server_t *s, uint64_t *jobs_held_for_budget,
                                         uint64_t *peak_used_bytes) {
    if (!s) return rpx_set_error(RPX_E_INVALID, "server is null");
    if (jobs_held_for_budget) *jobs_held_for_budget = s->jobs_held_for_budget;
    if (peak_used_bytes) *peak_used_bytes = s->peak_used;
    return RPX_OK;
}

extern "C" rpx_status_t rpx_server_counts(const rpx_server_t *s, uint32_t *waiting, uint32_t *active) {
    if (!s) return rpx_set_error(RPX_E_INVALID, "server is null");
    if (waiting) *waiting = (uint32_t)s->waiting.size();
    if (active) *active = (uint32_t)s->active.size();
    return RPX_OK;
}

extern "C" rpx_status_t rpx_server_set_graph_launch(rpx_server_t *s, int32_t on) {
    if (!s) return rpx_set_error(RPX_E_INVALID, "null server");
    s->graph_launch = on != 0;
    return RPX_OK;
}
