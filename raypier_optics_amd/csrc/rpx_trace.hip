// Generation tracer for gfx950: one generation of rays (or gausslets) against
// a device-resident scene, with stable parent-major compaction of children.
//
// One launch per generation (after k_clear): each workgroup takes a tile of
// 256 rays, runs the nearest-face search (scene staged in LDS), updates the
// parents in place (length, end_face_idx), counts face hits, evaluates the
// material, and places the children by a decoupled look-back over the tiles
// in parent-major order (k_generation for rays; k_gausslet_generation for
// gausslets, which builds children and parabasal rays after the look-back).
// rpx_trace_rays chains the ray generations on the device (rpx_gen.h).
#include <hip/hip_runtime.h>
#include <atomic>
#include <algorithm>
#include <string.h>
#include <stdio.h>
#include <string>
#include <vector>
#include <mutex>
#include "rpx_materials.h"
#include "rpx_host.h"
#include "rpx_heap.h"
#include <map>
#include "rpx_gen.h"

using namespace rpx;

// ---------------------------------------------------------------- errors
static thread_local std::string g_last_error;

rpx_status_t rpx_set_error(rpx_status_t code, const std::string &msg) {
    g_last_error = msg;
    return code;
}
static rpx_status_t set_error(rpx_status_t code, const std::string &msg) { return rpx_set_error(code, msg); }
#define HIP_TRY(expr) RPX_HIP_TRY(expr)

extern "C" const char *rpx_last_error(void) { return g_last_error.c_str(); }
extern "C" int32_t rpx_abi_version(void) { return RPX_ABI_VERSION; }

extern "C" int32_t rpx_struct_size(const char *name) {
    if (!name) return -1;
    std::string n(name);
    if (n == "rpx_transform_t") return sizeof(rpx_transform_t);
    if (n == "rpx_face_list_t") return sizeof(rpx_face_list_t);
    if (n == "rpx_face_t") return sizeof(rpx_face_t);
    if (n == "rpx_shape_op_t") return sizeof(rpx_shape_op_t);
    if (n == "rpx_dispersion_t") return sizeof(rpx_dispersion_t);
    if (n == "rpx_material_t") return sizeof(rpx_material_t);
    if (n == "rpx_zernike_insn_t") return sizeof(rpx_zernike_insn_t);
    if (n == "rpx_zernike_coef_t") return sizeof(rpx_zernike_coef_t);
    if (n == "rpx_zernike_t") return sizeof(rpx_zernike_t);
    if (n == "rpx_rays_t") return sizeof(rpx_rays_t);
    if (n == "rpx_generation_t") return sizeof(rpx_generation_t);
    if (n == "rpx_lattice_t") return sizeof(rpx_lattice_t);
    if (n == "rpx_scene_desc_t") return sizeof(rpx_scene_desc_t);
    if (n == "rpx_rays_split_t") return sizeof(rpx_rays_split_t);
    if (n == "rpx_job_desc_t") return sizeof(rpx_job_desc_t);
    if (n == "rpx_job_result_t") return sizeof(rpx_job_result_t);
    if (n == "rpx_lens_t") return sizeof(rpx_lens_t);
    return -1;
}

// ----------------------------------------------------------------- scene

static int material_max_children(const rpx_material_t &m) {
    switch (m.type) {
    case RPX_MAT_GRATING:
        return m.n_orders > 1 ? m.n_orders : 1;
    case RPX_MAT_TRANSPARENT:
    case RPX_MAT_PEC:
    case RPX_MAT_DIELECTRIC:
        return 1;
    case RPX_MAT_PARTIAL:
    case RPX_MAT_FULL_DIELECTRIC:
    case RPX_MAT_FULL_DISPERSIVE:
    case RPX_MAT_SINGLE_LAYER:
    case RPX_MAT_COATED_DISPERSIVE:
        return 2;
    default:
        return 0;
    }
}

rpx_status_t rpx_check_shape_program(const rpx_shape_op_t *ops, int n_ops, int root) {
    if (root < 0) return RPX_OK;
    if (root >= n_ops) return set_error(RPX_E_INVALID, "shape root out of range");
    int begin = ops[root].begin, depth = 0;
    if (begin < 0 || begin > root) return set_error(RPX_E_INVALID, "shape program range");
    for (int k = begin; k <= root; ++k) {
        int op = ops[k].op;
        if (op < RPX_SHAPE_TRUE || op > RPX_SHAPE_XOR) return set_error(RPX_E_INVALID, "shape opcode");
        if (op <= RPX_SHAPE_POLYGON) ++depth;
        else if (op == RPX_SHAPE_NOT) { if (depth < 1) return set_error(RPX_E_INVALID, "shape stack"); }
        else { if (depth < 2) return set_error(RPX_E_INVALID, "shape stack"); --depth; }
        if (depth > 31) return set_error(RPX_E_UNSUPPORTED, "shape program deeper than 31");
    }
    if (depth != 1) return set_error(RPX_E_INVALID, "shape program does not reduce to one value");
    return RPX_OK;
}

static rpx_status_t validate_scene(const rpx_scene_desc_t *d) {
    if (!d) return set_error(RPX_E_INVALID, "scene desc is null");
    if (d->n_face_lists < 0 || d->n_faces < 0 || d->n_materials < 0)
        return set_error(RPX_E_INVALID, "negative table size");
    if (d->n_faces > RPX_LDS_FACES)
        return set_error(RPX_E_UNSUPPORTED, "more than 64 faces in one scene");
    if (d->n_face_lists > RPX_LDS_LISTS)
        return set_error(RPX_E_UNSUPPORTED, "more than 16 face lists in one scene");
    int expect = 0;
    for (int l = 0; l < d->n_face_lists; ++l) {
        const rpx_face_list_t &fl = d->face_lists[l];
        if (fl.first_face != expect || fl.n_faces < 0)
            return set_error(RPX_E_INVALID, "face lists must tile the face table in order");
        expect += fl.n_faces;
    }
    if (expect > d->n_faces) return set_error(RPX_E_INVALID, "face lists exceed face table");
    for (int l = 0; l < d->n_face_lists; ++l)
        for (int k = d->face_lists[l].first_face; k < d->face_lists[l].first_face + d->face_lists[l].n_faces; ++k)
            if (d->faces[k].face_list != l) return set_error(RPX_E_INVALID, "face_list back-reference");
    for (int i = 0; i < d->n_faces; ++i) {
        const rpx_face_t &f = d->faces[i];
        if (f.type < RPX_FACE_CIRCULAR || f.type > RPX_FACE_DISTORTION)
            return set_error(RPX_E_UNSUPPORTED, "unknown face type " + std::to_string(f.type));
        if (f.material >= d->n_materials) return set_error(RPX_E_INVALID, "face material index");
        if (f.material >= 0 && d->materials[f.material].type == RPX_MAT_GRATING &&
            (d->materials[f.material].n_orders < 0 || d->materials[f.material].n_orders > 3))
            return set_error(RPX_E_INVALID, "a grating has more than three orders");
        if (f.shape >= d->n_shape_ops) return set_error(RPX_E_INVALID, "face shape index");
        if (f.type == RPX_FACE_POLYGON &&
            (f.aux1 < 1 || f.aux0 < 0 || f.aux0 + f.aux1 > d->n_vertices))
            return set_error(RPX_E_INVALID, "polygon vertex range");
        if (f.type == RPX_FACE_DISTORTION) {
            if (f.aux0 < 0 || f.aux0 >= d->n_faces || d->faces[f.aux0].type == RPX_FACE_DISTORTION)
                return set_error(RPX_E_INVALID, "distortion base face");
            if (f.aux1 < 0 || f.aux1 >= d->n_zernikes) return set_error(RPX_E_INVALID, "distortion index");
        }
    }
    for (int i = 0; i < d->n_shape_ops; ++i) {
        const rpx_shape_op_t &o = d->shape_ops[i];
        if (o.op < RPX_SHAPE_TRUE || o.op > RPX_SHAPE_XOR) return set_error(RPX_E_INVALID, "shape opcode");
        if (o.op == RPX_SHAPE_POLYGON && (o.b < 1 || o.a < 0 || o.a + o.b > d->n_vertices))
            return set_error(RPX_E_INVALID, "shape polygon range");
    }
    for (int i = 0; i < d->n_faces; ++i) {  // postfix programs must be well formed
        rpx_status_t st = rpx_check_shape_program(d->shape_ops, d->n_shape_ops, d->faces[i].shape);
        if (st != RPX_OK) return st;
    }
    for (int i = 0; i < d->n_materials; ++i) {
        const rpx_material_t &m = d->materials[i];
        if (m.type < RPX_MAT_NONE || m.type > RPX_MAT_GRATING)
            return set_error(RPX_E_UNSUPPORTED, "unknown material type " + std::to_string(m.type));
        const rpx_dispersion_t *ds[3] = {&m.disp_inside, &m.disp_outside, &m.disp_coating};
        for (auto *dd : ds)
            if (dd->n_coefs < 0 || dd->n_coefs > RPX_MAX_DISP_COEFS)
                return set_error(RPX_E_INVALID, "dispersion coefficient count");
    }
    for (int i = 0; i < d->n_zernikes; ++i) {
        const rpx_zernike_t &z = d->zernikes[i];
        if (z.n_slots < 0 || z.n_slots > RPX_MAX_ZERNIKE_SLOTS)
            return set_error(RPX_E_UNSUPPORTED, "Zernike workspace too large");
        if (z.coef_begin < 0 || z.coef_begin + z.n_coefs > d->n_zernike_coefs)
            return set_error(RPX_E_INVALID, "Zernike coefficient range");
        if (z.sag_begin < 0 || z.sag_begin + z.sag_len > d->n_zernike_insns ||
            z.grad_begin < 0 || z.grad_begin + z.grad_len > d->n_zernike_insns)
            return set_error(RPX_E_INVALID, "Zernike program range");
        for (int k = 0; k < z.n_coefs; ++k) {
            const rpx_zernike_coef_t &c = d->zernike_coefs[z.coef_begin + k];
            int refs[4] = {c.r_ref, c.rp_ref, c.ror_ref, c.sag_ref};
            for (int r : refs)
                if (r < -2 || r >= z.n_slots) return set_error(RPX_E_INVALID, "Zernike value ref");
            for (int q = 0; q < 4; ++q)
                if (c.deg[q] < 0 || c.deg[q] >= RPX_ZPOLY) return set_error(RPX_E_INVALID, "Zernike polynomial degree");
        }
        int ranges[2][2] = {{z.sag_begin, z.sag_len}, {z.grad_begin, z.grad_len}};
        for (auto &rg : ranges)
            for (int k = rg[0]; k < rg[0] + rg[1]; ++k) {
                const rpx_zernike_insn_t &in = d->zernike_insns[k];
                if (in.dst < 0 || in.dst >= z.n_slots || in.op < 0 || in.op > 2)
                    return set_error(RPX_E_INVALID, "Zernike instruction");
                for (int s = 0; s < 5; ++s)
                    if (in.src[s] < -2 || in.src[s] >= z.n_slots)
                        return set_error(RPX_E_INVALID, "Zernike instruction operand");
            }
    }
    return RPX_OK;
}

template <class T>
static size_t push_table(std::vector<char> &blob, const T *src, int n) {
    size_t off = (blob.size() + 255) & ~size_t(255);
    blob.resize(off + sizeof(T) * (size_t)(n > 0 ? n : 0));
    if (n > 0) memcpy(blob.data() + off, src, sizeof(T) * (size_t)n);
    return off;
}

// The device tables of every scene on one device come from one heap: 2 MiB
// chunks carved into blocks, so creating and destroying a scene makes no
// hipMalloc / hipFree (hipFree waits for the whole device: a serving job's
// submission paid about 0.5 ms for it, up to 55 ms).  A destroyed scene's
// block is reused by the next scene at once; no work that reads a scene may
// be in flight when it is destroyed.  The heaps live until the process ends.
namespace {
struct SceneHeaps {
    std::mutex m;
    std::map<int, rpx::BlockHeap<DeviceChunks> *> by_device;
};

SceneHeaps &scene_heaps() {
    static SceneHeaps *h = new SceneHeaps();
    return *h;
}

void *scene_block_alloc(int dev, uint64_t bytes) {
    SceneHeaps &H = scene_heaps();
    std::lock_guard<std::mutex> lock(H.m);
    rpx::BlockHeap<DeviceChunks> *&heap = H.by_device[dev];
    if (!heap) {
        heap = new rpx::BlockHeap<DeviceChunks>();
        heap->set_limits(64ull << 20, 2ull << 20);
    }
    return heap->alloc(bytes);
}

void scene_block_free(int dev, void *p) {
    SceneHeaps &H = scene_heaps();
    std::lock_guard<std::mutex> lock(H.m);
    auto it = H.by_device.find(dev);
    if (it != H.by_device.end() && it->second) (void)it->second->free(p);
}
}  // namespace

extern "C" rpx_status_t rpx_scene_create(const rpx_scene_desc_t *d, rpx_scene_t **out) {
    if (!out) return set_error(RPX_E_INVALID, "out is null");
    *out = nullptr;
    rpx_status_t st = validate_scene(d);
    if (st != RPX_OK) return st;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_error(RPX_E_NO_DEVICE, "no HIP device");
    std::vector<char> blob;
    size_t o_fl = push_table(blob, d->face_lists, d->n_face_lists);
    size_t o_f = push_table(blob, d->faces, d->n_faces);
    size_t o_m = push_table(blob, d->materials, d->n_materials);
    size_t o_s = push_table(blob, d->shape_ops, d->n_shape_ops);
    size_t o_v = push_table(blob, d->vertices, 2 * d->n_vertices);
    size_t o_z = push_table(blob, d->zernikes, d->n_zernikes);
    size_t o_zc = push_table(blob, d->zernike_coefs, d->n_zernike_coefs);
    size_t o_zi = push_table(blob, d->zernike_insns, d->n_zernike_insns);
    blob.resize(((blob.size() + 255) & ~size_t(255)) + 256);
    int device = 0;
    HIP_TRY(hipGetDevice(&device));
    char *dev = (char *)scene_block_alloc(device, blob.size());
    if (!dev) return set_error(RPX_E_HIP, "scene upload: out of device memory");
    hipError_t e = hipMemcpy(dev, blob.data(), blob.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        scene_block_free(device, dev);
        return set_error(RPX_E_HIP, std::string("scene upload: ") + hipGetErrorString(e));
    }
    rpx_scene *s = new rpx_scene();
    s->device_block = dev;
    s->device = device;
    SceneView &v = s->view;
    v.face_lists = (const rpx_face_list_t *)(dev + o_fl);
    v.faces = (const rpx_face_t *)(dev + o_f);
    v.materials = (const rpx_material_t *)(dev + o_m);
    v.shape_ops = (const rpx_shape_op_t *)(dev + o_s);
    v.vertices = (const double *)(dev + o_v);
    v.zernikes = (const rpx_zernike_t *)(dev + o_z);
    v.zcoefs = (const rpx_zernike_coef_t *)(dev + o_zc);
    v.zinsns = (const rpx_zernike_insn_t *)(dev + o_zi);
    v.n_face_lists = d->n_face_lists;
    v.n_faces = d->n_faces;
    v.n_materials = d->n_materials;
    v.n_shape_ops = d->n_shape_ops;
    v.n_vertices = d->n_vertices;
    v.n_zernikes = d->n_zernikes;
    v.n_zcoefs = d->n_zernike_coefs;
    v.n_zinsns = d->n_zernike_insns;
    int nlist = 0, maxc = 0, nglob = 0;
    for (int l = 0; l < d->n_face_lists; ++l) nlist += d->face_lists[l].n_faces;
    v.n_list_faces = nlist;
    for (int i = 0; i < nlist; ++i) {
        const rpx_face_t &f = d->faces[i];
        if (f.global_idx + 1 > nglob) nglob = f.global_idx + 1;
        int mc = f.material >= 0 ? material_max_children(d->materials[f.material]) : 0;
        if (mc > maxc) maxc = mc;
    }
    v.max_children = maxc;
    uint32_t feat = 0;
    for (int i = 0; i < d->n_faces; ++i) {
        switch (d->faces[i].type) {
        case RPX_FACE_SPHERICAL: case RPX_FACE_SHAPED_SPHERICAL: feat |= F_SPHERE; break;
        case RPX_FACE_CONIC: feat |= F_CONIC; break;
        case RPX_FACE_ASPHERIC: feat |= F_ASPHERE; break;
        case RPX_FACE_DISTORTION: feat |= F_DISTORT; break;
        default: feat |= F_PLANE; break;
        }
    }
    for (int i = 0; i < d->n_materials; ++i) {
        switch (d->materials[i].type) {
        case RPX_MAT_FULL_DIELECTRIC: case RPX_MAT_FULL_DISPERSIVE: feat |= M_FRESNEL; break;
        case RPX_MAT_SINGLE_LAYER: case RPX_MAT_COATED_DISPERSIVE: feat |= M_COATED; break;
        case RPX_MAT_GRATING: feat |= M_GRATING; break;
        default: feat |= M_SIMPLE; break;
        }
    }
    s->features = feat;
    s->n_global_faces = nglob;
    s->negative_face_idx = false;
    for (int i = 0; i < nlist; ++i)
        if (d->faces[i].global_idx < 0) s->negative_face_idx = true;
    s->host_shape_ops.assign(d->shape_ops, d->shape_ops + d->n_shape_ops);
    s->face_global_idx.resize(d->n_faces);
    for (int i = 0; i < d->n_faces; ++i) s->face_global_idx[i] = d->faces[i].global_idx;
    for (int l = 0; l < d->n_face_lists; ++l) s->list_sizes[l] = d->face_lists[l].n_faces;
    *out = s;
    return RPX_OK;
}

extern "C" rpx_status_t rpx_scene_destroy(rpx_scene_t *s) {
    if (!s) return RPX_OK;
    scene_block_free(s->device, s->device_block);
    delete s;
    return RPX_OK;
}

extern "C" int32_t rpx_scene_max_children(const rpx_scene_t *s) { return s ? s->view.max_children : -1; }

extern "C" uint64_t rpx_trace_workspace_bytes(uint64_t n_in) {
    uint64_t nb = (n_in + RPX_BLOCK - 1) / RPX_BLOCK;
    return align256(4 * n_in) + 2 * align256(4 * nb) + 2 * align256(8 * nb) + align256(sizeof(GenControl));
}

// --------------------------------------------------------------- kernels
struct TraceArgs {
    SceneView S;
    rpx_rays_t in;
    rpx_rays_t out;
    uint64_t n_in;
    double max_length;      // already rounded to float for ray generations
    WlTable wavelengths;
    uint32_t *face_hits;
    Workspace ws;
    int32_t one_list, one_face;  // one-face mode: list and face index inside it (-1 = off)
    // device-resident generation chain (k_generation only; link = nullptr
    // otherwise): this generation's count and offsets are read from link[0]
    // and its children's written to link[1]; `in` (when in_pool) and `out`
    // are the pool with its field stride, the children bounded by
    // pool_rays - link[0].out_offset, or not stored when !store_children
    GenLink *link;
    uint64_t pool_rays;
    int32_t in_pool, store_children;
    uint64_t status_tag;     // look-back tag of this generation (0 outside a chain)
    GenControl *head_out;    // chain: where the last workgroup to finish writes the control head
    // tile order (k_generation): with by_launch, tile = tile_base +
    // blockIdx.x and look-back waits are bounded by lookback_ticks (ChainGen);
    // otherwise tiles come from the generation's atomic counter
    uint32_t tile_base = 0;
    int32_t by_launch = 0;
    uint64_t lookback_ticks = 0;
};

struct CountEmit {
    static constexpr int kCap = 3;
    int n = 0;
    RPX_DEV void operator()(const Ray &) { ++n; }
    RPX_DEV void skip() { ++n; }
};

// Writes at most `room` children (the count reserved by the counting pass),
// so a lane can never write outside its slot range.
struct StoreEmit {
    static constexpr int kCap = 3;
    const rpx_rays_t *out;
    uint64_t pos;
    uint32_t room;
    RPX_DEV void operator()(const Ray &r) {
        if (room == 0) return;
        store_ray(*out, pos++, r);
        --room;
    }
    RPX_DEV void skip() {}
};

// After a generation's children are placed: stores children from index 2 on
// (a multi-order grating's third order) at pos + index, within `limit`.
struct TailEmit {
    static constexpr int kCap = 3;
    const rpx_rays_t *out;
    uint64_t pos, limit;
    int n = 0;
    RPX_DEV void operator()(const Ray &r) {
        if (n >= 2 && pos + n < limit) store_ray(*out, pos + n, r);
        ++n;
    }
    RPX_DEV void skip() { ++n; }
};

// Collects up to 3 children of one gausslet base ray.
struct CollectEmit {
    static constexpr int kCap = 3;
    Ray r[3];
    int n = 0;
    RPX_DEV void operator()(const Ray &x) {
        if (n < 3) r[n] = x;
        ++n;
    }
    RPX_DEV void skip() { ++n; }
};

// One gausslet generation in one launch (trace_gausslet_c, ctracer.pyx:
// 2214-2281, with trace_parabasal_rays, :2350-2385): each lane takes one
// gausslet; the base ray's nearest-face search and its parabasal rays'
// intersections with that face (every one must hit it, lengths updated in
// place) come first, with the child count from the material; the decoupled
// look-back then places the children, and the lane builds them and their
// parabasal rays straight into their slots (children's lengths reset to
// max_length, :2280).  Gausslet children are built after the look-back, not
// held in registers through it: a child carries six parabasal rays.
template <uint32_t F>
__global__ __launch_bounds__(RPX_BLOCK) void k_gausslet_generation(TraceArgs A) {
    __shared__ SceneLds lds_scene;
    __shared__ uint32_t lds_hist[RPX_LDS_FACES];
    __shared__ uint32_t lds_wave[RPX_BLOCK / 64];
    __shared__ uint32_t lds_tile, lds_hits, lds_parents;
    __shared__ uint64_t lds_excl;
    // in a device-resident chain, size and pool place come from the link,
    // and tiles follow the chain's tile order (as k_generation)
    uint64_t n_in = A.n_in, out_limit = A.out.capacity;
    rpx_rays_t in = A.in, out = A.out;
    if (A.link) {
        const GenLink L = A.link[0];
        n_in = L.count;
        if (A.in_pool) in = advance_rays(A.in, L.offset);
        out = advance_rays(A.out, L.out_offset);
        out_limit = A.store_children ? (A.pool_rays > L.out_offset ? A.pool_rays - L.out_offset : 0) : 0;
    }
    const uint64_t n_tiles = (n_in + RPX_BLOCK - 1) / RPX_BLOCK;
    uint32_t my_tile = 0;
    if (A.by_launch) {
        my_tile = A.tile_base + blockIdx.x;
        if (my_tile >= n_tiles) return;
    } else {
        if (blockIdx.x >= n_tiles) return;
        if (threadIdx.x == 0) my_tile = atomicAdd(&A.ws.ctl->tile_counter, 1u);
    }
    const SceneView S = stage_scene(A.S, lds_scene);
    for (int k = threadIdx.x; k < RPX_LDS_FACES; k += blockDim.x) lds_hist[k] = 0;
    if (threadIdx.x == 0) {
        lds_tile = my_tile;
        lds_hits = 0;
        lds_parents = 0;
    }
    __syncthreads();
    const uint32_t tile = lds_tile;
    const uint64_t i = (uint64_t)tile * RPX_BLOCK + threadIdx.x;
    const uint64_t c = in.capacity;
    uint32_t count = 0;
    int hit_rec = -1, hit_list = -1;
    if (i < n_in) {
        const double *f = in.f64;
        V3 origin = v3(f[RPX_OX * c + i], f[RPX_OY * c + i], f[RPX_OZ * c + i]);
        V3 dir = v3(f[RPX_DX * c + i], f[RPX_DY * c + i], f[RPX_DZ * c + i]);
        // in a chain every gausslet starts at max_length (trace_rays resets
        // the input, children are written so), which keeps a generation
        // repeatable after a run-out look-back (its in-place lengths are then
        // not read back)
        double length = A.link ? A.max_length : f[RPX_LENGTH * c + i];
        uint32_t end_face = RPX_NO_FACE;
        V3 point = add(origin, mul(dir, A.max_length));
        if (A.one_list < 0) {
            // trace_segment_c / FaceList.intersect_c (ctracer.pyx:2093-2104, 1882-1904)
            for (int l = 0; l < S.n_face_lists; ++l) {
                const rpx_face_list_t &fl = S.face_lists[l];
                V3 p1 = transform(fl.inv_trans, origin);
                V3 p2 = transform(fl.inv_trans, point);
                const double slen = sep(p1, p2), zrate = (p2.z - p1.z) / slen;
                for (int k = fl.first_face; k < fl.first_face + fl.n_faces; ++k) {
                    const rpx_face_t &fc = S.faces[k];
                    if (face_out_of_reach(fc, p1.z, zrate * length, slen)) continue;
                    double dist = face_intersect<F>(S, fc, p1, p2, 1);
                    if (fc.tolerance < dist && dist < length) {
                        length = dist;
                        end_face = (uint32_t)fc.global_idx;
                        hit_rec = k;
                        hit_list = l;
                    }
                }
            }
        } else {
            // FaceList.intersect_one_face_c (ctracer.pyx:1861-1879)
            const rpx_face_list_t &fl = S.face_lists[A.one_list];
            const int k = fl.first_face + A.one_face;
            const rpx_face_t &fc = S.faces[k];
            V3 p1 = transform(fl.inv_trans, origin);
            V3 p2 = transform(fl.inv_trans, point);
            double dist = face_intersect<F>(S, fc, p1, p2, 1);
            if (fc.tolerance < dist && dist < length) {
                length = dist;
                end_face = (uint32_t)fc.global_idx;
                hit_rec = k;
                hit_list = A.one_list;
            }
        }
        in.f64[RPX_LENGTH * c + i] = length;
        in.u32[RPX_END_FACE_IDX * c + i] = end_face;
        if (hit_rec >= 0) {
            atomicAdd(&lds_hist[hit_rec], 1u);
            atomicAdd(&lds_hits, 1u);
            const rpx_face_t &fc = S.faces[hit_rec];
            const rpx_face_list_t &fl = S.face_lists[hit_list];
            // trace_parabasal_rays (ctracer.pyx:2363-2373): every parabasal
            // ray must hit the same face; lengths update in place
            bool para_ok = true;
            const double *P = in.para;
            for (int j = 0; j < RPX_N_PARA && para_ok; ++j) {
                const uint64_t b = (uint64_t)j * RPX_PARA_FIELDS;
                V3 po = v3(P[(b + 0) * c + i], P[(b + 1) * c + i], P[(b + 2) * c + i]);
                V3 pd = v3(P[(b + 3) * c + i], P[(b + 4) * c + i], P[(b + 5) * c + i]);
                double plen = A.link ? A.max_length : P[(b + 9) * c + i];
                V3 pend = add(po, mul(pd, A.max_length));
                V3 q1 = transform(fl.inv_trans, po);
                V3 q2 = transform(fl.inv_trans, pend);
                double dist = face_intersect<F>(S, fc, q1, q2, 0);
                if (fc.tolerance < dist && dist < plen) {
                    in.para[(b + 9) * c + i] = dist;
                } else {
                    para_ok = false;
                }
            }
            if (para_ok && fc.material >= 0) {
                Ray r = load_ray(in, i);
                r.length = length;
                r.end_face_idx = end_face;
                V3 hp = add(r.origin, mul(r.direction, length));
                Orientation o = compute_orientation<F>(S, fl, fc, hp);
                CountEmit ce;
                eval_children<F>(S.materials[fc.material], r, (uint32_t)i, hp, o, A.wavelengths, ce);
                count = (uint32_t)min(ce.n, 3);
            }
        }
    }
    uint32_t block_total;
    const uint32_t local = block_scan_small(count, lds_wave, &block_total);
    {
        const uint64_t mp = __ballot(count > 0);
        if ((threadIdx.x & 63) == 0 && mp) atomicAdd(&lds_parents, (uint32_t)__popcll(mp));
    }
    if (threadIdx.x < 64) {
        const uint64_t excl = lookback_tagged(A.ws.tile_status, tile, 0, block_total, &A.ws.ctl->error, A.status_tag,
                                              A.lookback_ticks);
        if (threadIdx.x == 0) lds_excl = excl;
    }
    __syncthreads();
    const uint64_t pos = lds_excl + local;
    if (count > 0 && pos + count <= out_limit) {
        const rpx_face_t &fc = S.faces[hit_rec];
        const rpx_face_list_t &fl = S.face_lists[hit_list];
        const rpx_material_t &m = S.materials[fc.material];
        Ray r = load_ray(in, i);
        V3 hp = add(r.origin, mul(r.direction, r.length));
        Orientation o = compute_orientation<F>(S, fl, fc, hp);
        CollectEmit ce;
        eval_children<F>(m, r, (uint32_t)i, hp, o, A.wavelengths, ce);
        const uint64_t co = out.capacity;
        const int nk = min(min(ce.n, 3), (int)count);
        for (int k = 0; k < nk; ++k) {
            Ray child = ce.r[k];
            child.length = A.max_length;
            store_ray(out, pos + k, child);
        }
        for (int j = 0; j < RPX_N_PARA; ++j) {
            const uint64_t b = (uint64_t)j * RPX_PARA_FIELDS;
            const double *P = in.para;
            V3 po = v3(P[(b + 0) * c + i], P[(b + 1) * c + i], P[(b + 2) * c + i]);
            V3 pd = v3(P[(b + 3) * c + i], P[(b + 4) * c + i], P[(b + 5) * c + i]);
            double plen = P[(b + 9) * c + i];
            V3 pp = add(po, mul(pd, plen));
            Orientation po_ = compute_orientation<F>(S, fl, fc, pp);
            for (int k = 0; k < nk; ++k) {
                Para q = eval_parabasal<F>(m, ce.r[k], pd, pp, po_, ce.r[k].ray_type_id, A.wavelengths);
                double *Q = out.para;
                const uint64_t oi = pos + k;
                Q[(b + 0) * co + oi] = q.origin.x;
                Q[(b + 1) * co + oi] = q.origin.y;
                Q[(b + 2) * co + oi] = q.origin.z;
                Q[(b + 3) * co + oi] = q.direction.x;
                Q[(b + 4) * co + oi] = q.direction.y;
                Q[(b + 5) * co + oi] = q.direction.z;
                Q[(b + 6) * co + oi] = q.normal.x;
                Q[(b + 7) * co + oi] = q.normal.y;
                Q[(b + 8) * co + oi] = q.normal.z;
                Q[(b + 9) * co + oi] = A.max_length;
            }
        }
    }
    if (threadIdx.x == 0) {
        add_hits(A.ws.ctl, tile, lds_hits);
        if (lds_parents) atomicAdd(&A.ws.ctl->parent_shards[tile % RPX_SHARDS], lds_parents);
        if ((uint64_t)tile == (n_in - 1) / RPX_BLOCK) {
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
    uint32_t *shard = A.ws.ctl->face_shards[tile % RPX_SHARDS];
    for (int k = threadIdx.x; k < S.n_list_faces; k += blockDim.x) {
        uint32_t h = lds_hist[k];
        if (h) atomicAdd(&shard[k], h);
    }
}

template <uint32_t F>
__global__ __launch_bounds__(RPX_BLOCK) __attribute__((amdgpu_waves_per_eu(F == V_ALL ? 1 : 3)))
void k_generation(TraceArgs A) {
    __shared__ SceneLds lds_scene;
    __shared__ uint32_t lds_hist[RPX_LDS_FACES];
    __shared__ uint32_t lds_wave[RPX_BLOCK / 64];
    __shared__ uint32_t lds_tile, lds_hits;
    __shared__ uint64_t lds_excl;
    // A generation of a device-resident chain learns its size and place in
    // the pool from the link its parent generation wrote.  Its grid is sized
    // on the host from a guess made before that size is known: workgroups
    // past the generation's tiles leave before taking a tile number, and
    // tiles past the grid are taken by a second launch of the same
    // generation (the tile counter and look-back carry on across it).
    uint64_t n_in = A.n_in, out_limit = A.out.capacity;
    rpx_rays_t in = A.in, out = A.out;
    if (A.link) {
        const GenLink L = A.link[0];
        n_in = L.count;
        if (A.in_pool) in = advance_rays(A.in, L.offset);
        out = advance_rays(A.out, L.out_offset);
        out_limit = A.store_children ? (A.pool_rays > L.out_offset ? A.pool_rays - L.out_offset : 0) : 0;
    }
    const uint32_t n_tiles = (uint32_t)((n_in + RPX_BLOCK - 1) / RPX_BLOCK);
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
    const double *f = in.f64;
    const uint64_t c = in.capacity;
    V3 origin = v3(0, 0, 0), dir = v3(0, 0, 0);
    if (A.by_launch) {
        const uint64_t i0 = (uint64_t)my_tile * RPX_BLOCK + threadIdx.x;
        if (i0 < n_in) {
            origin = v3(f[RPX_OX * c + i0], f[RPX_OY * c + i0], f[RPX_OZ * c + i0]);
            dir = v3(f[RPX_DX * c + i0], f[RPX_DY * c + i0], f[RPX_DZ * c + i0]);
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
    const uint64_t i = (uint64_t)tile * RPX_BLOCK + threadIdx.x;
    PairEmit ce;
    bool hit = false;
    int hit_rec = -1, hit_list = -1;
    if (i < n_in) {
        if (!A.by_launch) {
            origin = v3(f[RPX_OX * c + i], f[RPX_OY * c + i], f[RPX_OZ * c + i]);
            dir = v3(f[RPX_DX * c + i], f[RPX_DY * c + i], f[RPX_DZ * c + i]);
        }
        double length = A.max_length;
        uint32_t end_face = RPX_NO_FACE;
        V3 point = add(origin, mul(dir, A.max_length));
        if (A.one_list < 0) {
            // trace_segment_c / FaceList.intersect_c (ctracer.pyx:2093-2104, 1882-1904)
            for (int l = 0; l < S.n_face_lists; ++l) {
                const rpx_face_list_t &fl = S.face_lists[l];
                V3 p1 = transform(fl.inv_trans, origin);
                V3 p2 = transform(fl.inv_trans, point);
                const double slen = sep(p1, p2), zrate = (p2.z - p1.z) / slen;
                for (int k = fl.first_face; k < fl.first_face + fl.n_faces; ++k) {
                    const rpx_face_t &fc = S.faces[k];
                    if (face_out_of_reach(fc, p1.z, zrate * length, slen)) continue;
                    double dist = face_intersect<F>(S, fc, p1, p2, 1);
                    if (fc.tolerance < dist && dist < length) {
                        length = dist;
                        end_face = (uint32_t)fc.global_idx;
                        hit_rec = k;
                        hit_list = l;
                    }
                }
            }
        } else {
            // FaceList.intersect_one_face_c (ctracer.pyx:1861-1879)
            const rpx_face_list_t &fl = S.face_lists[A.one_list];
            const int k = fl.first_face + A.one_face;
            const rpx_face_t &fc = S.faces[k];
            V3 p1 = transform(fl.inv_trans, origin);
            V3 p2 = transform(fl.inv_trans, point);
            double dist = face_intersect<F>(S, fc, p1, p2, 1);
            if (fc.tolerance < dist && dist < length) {
                length = dist;
                end_face = (uint32_t)fc.global_idx;
                hit_rec = k;
                hit_list = A.one_list;
            }
        }
        in.f64[RPX_LENGTH * c + i] = length;
        in.u32[RPX_END_FACE_IDX * c + i] = end_face;
        if (hit_rec >= 0) {
            atomicAdd(&lds_hist[hit_rec], 1u);
            const rpx_face_t &fc = S.faces[hit_rec];
            if (fc.material >= 0 && material_emits(S.materials[fc.material])) {
                hit = true;   // counted: the rest of its record is read
                const rpx_face_list_t &fl = S.face_lists[hit_list];
                Ray r = load_ray(in, i);
                r.length = length;
                r.end_face_idx = end_face;
                V3 hp = add(r.origin, mul(r.direction, length));
                Orientation o = compute_orientation<F>(S, fl, fc, hp);
                eval_children<F>(S.materials[fc.material], r, (uint32_t)i, hp, o, A.wavelengths, ce);
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
    if (count > 0 && pos < out_limit) store_ray(out, pos, ce.first());
    if (count > 1 && pos + 1 < out_limit) store_ray(out, pos + 1, ce.r0);
    if constexpr ((F & M_GRATING) != 0) {
        if (count > 2 && pos + 2 < out_limit) {
            // the third child is built again now that its position is known
            const rpx_face_t &fc = S.faces[hit_rec];
            Ray r = load_ray(in, i);
            V3 hp = add(r.origin, mul(r.direction, r.length));
            Orientation o = compute_orientation<F>(S, S.face_lists[hit_list], fc, hp);
            TailEmit te{&out, pos, out_limit};
            grating_children(S.materials[fc.material], r, (uint32_t)i, hp, o, A.wavelengths, te);
        }
    }
    if (threadIdx.x == 0) {
        // a chain's hit total is the sum of its face totals (k_chain_fold);
        // one word per generation would take every tile's atomic in one line
        if (!A.link) add_hits(A.ws.ctl, tile, lds_hits);
        if ((uint64_t)tile == (n_in - 1) / RPX_BLOCK) {
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

// ------------------------------------------------------------- variants
static void launch_generation(uint32_t variant, dim3 grid, hipStream_t stream, const TraceArgs &A) {
    if (variant == V_SPHERES) hipLaunchKernelGGL((k_generation<V_SPHERES>), grid, dim3(RPX_BLOCK), 0, stream, A);
    else if (variant == V_LENS) hipLaunchKernelGGL((k_generation<V_LENS>), grid, dim3(RPX_BLOCK), 0, stream, A);
    else hipLaunchKernelGGL((k_generation<V_ALL>), grid, dim3(RPX_BLOCK), 0, stream, A);
}

static void launch_gausslet_generation(uint32_t variant, dim3 grid, hipStream_t stream, const TraceArgs &A) {
    if (variant == V_SPHERES)
        hipLaunchKernelGGL((k_gausslet_generation<V_SPHERES>), grid, dim3(RPX_BLOCK), 0, stream, A);
    else if (variant == V_LENS)
        hipLaunchKernelGGL((k_gausslet_generation<V_LENS>), grid, dim3(RPX_BLOCK), 0, stream, A);
    else
        hipLaunchKernelGGL((k_gausslet_generation<V_ALL>), grid, dim3(RPX_BLOCK), 0, stream, A);
}

// ------------------------------------------------------------- timing
static std::mutex g_timing_mutex;
static bool g_timing_on = false;
static double g_timing_ms[RPX_TIMING_SLOTS] = {};
static uint64_t g_timing_launches[RPX_TIMING_SLOTS] = {};
static double g_timing_bytes[RPX_TIMING_SLOTS] = {};

bool rpx_timing_is_on() {
    std::lock_guard<std::mutex> g(g_timing_mutex);
    return g_timing_on;
}

void rpx_timing_add(int slot, float ms, double bytes) {
    if (slot < 0 || slot >= RPX_TIMING_SLOTS) return;
    std::lock_guard<std::mutex> g(g_timing_mutex);
    g_timing_ms[slot] += ms;
    g_timing_launches[slot] += 1;
    g_timing_bytes[slot] += bytes;
}

extern "C" rpx_status_t rpx_timing_enable(int32_t on) {
    std::lock_guard<std::mutex> g(g_timing_mutex);
    g_timing_on = on != 0;
    return RPX_OK;
}

extern "C" rpx_status_t rpx_timing_read(double *ms, uint64_t *launches, double *bytes) {
    std::lock_guard<std::mutex> g(g_timing_mutex);
    for (int k = 0; k < RPX_TIMING_SLOTS; ++k) {
        if (ms) ms[k] = g_timing_ms[k];
        if (launches) launches[k] = g_timing_launches[k];
        if (bytes) bytes[k] = g_timing_bytes[k];
        g_timing_ms[k] = 0;
        g_timing_launches[k] = 0;
        g_timing_bytes[k] = 0;
    }
    return RPX_OK;
}

// ------------------------------------------------ tile order of the chains
static std::atomic<uint64_t> g_lookback_ticks{0};
static std::atomic<uint64_t> g_tile_order_fallbacks{0};

uint64_t rpx_lookback_ticks_override() { return g_lookback_ticks.load(); }
void rpx_note_tile_order_fallback() { g_tile_order_fallbacks.fetch_add(1); }

extern "C" rpx_status_t rpx_debug_lookback_ticks(uint64_t ticks) {
    g_lookback_ticks.store(ticks);
    return RPX_OK;
}

extern "C" uint64_t rpx_tile_order_fallbacks(void) { return g_tile_order_fallbacks.load(); }

// ------------------------------------------------------------- host side
rpx_status_t rpx_check_trace_scene(const rpx_scene_t *scene) {
    if (!scene) return set_error(RPX_E_INVALID, "scene is null");
    if (scene->negative_face_idx)
        return set_error(RPX_E_INVALID, "a face has a negative global index (Face.idx): such a scene can be used "
                                        "for capture selection but not traced");
    return RPX_OK;
}

static rpx_status_t check_rays(const rpx_rays_t *r, uint64_t n, bool gauss, const char *what) {
    if (!r) return set_error(RPX_E_INVALID, std::string(what) + " is null");
    if (r->capacity < n) return set_error(RPX_E_INVALID, std::string(what) + " capacity below count");
    if (n > 0 && (!r->f64 || !r->u32)) return set_error(RPX_E_INVALID, std::string(what) + " has null fields");
    if (gauss && n > 0 && !r->para) return set_error(RPX_E_INVALID, std::string(what) + " lacks parabasal rays");
    return RPX_OK;
}

static rpx_status_t run_generation(const rpx_scene_t *scene, int one_list, int one_face, bool gauss,
                                   const rpx_rays_t *in, uint64_t n_in, double max_length,
                                   const double *wavelengths, uint64_t n_wavelengths, rpx_rays_t *out,
                                   uint64_t *n_out, uint32_t *face_hits, void *workspace,
                                   uint64_t workspace_bytes, void *stream_) {
    if (!n_out) return set_error(RPX_E_INVALID, "n_out is null");
    *n_out = 0;
    rpx_status_t st = rpx_check_trace_scene(scene);
    if (st) return st;
    st = check_rays(in, n_in, gauss, "input rays");
    if (st) return st;
    if (!out) return set_error(RPX_E_INVALID, "output rays are null");
    if (n_in == 0) return RPX_OK;
    if (n_in > 0xFFFFFFFFull) return set_error(RPX_E_UNSUPPORTED, "more than 2^32-1 rays in one generation");
    if (!wavelengths && n_wavelengths) return set_error(RPX_E_INVALID, "wavelength table is null");
    if (workspace_bytes < rpx_trace_workspace_bytes(n_in) || !workspace)
        return set_error(RPX_E_INVALID, "workspace too small");
    if (one_list >= 0) {
        if (one_list >= scene->view.n_face_lists || one_face < 0 || one_face >= scene->list_sizes[one_list])
            return set_error(RPX_E_INVALID, "face list / face index out of range");
    }
    hipStream_t stream = (hipStream_t)stream_;
    TraceArgs A;
    A.S = scene->view;
    A.in = *in;
    A.out = *out;
    A.n_in = n_in;
    A.max_length = max_length;
    A.wavelengths.wl = wavelengths;
    A.wavelengths.n = wavelengths ? n_wavelengths : 0;
    A.face_hits = face_hits;
    A.ws = carve(workspace, n_in);
    A.one_list = one_list;
    A.one_face = one_face;
    A.link = nullptr;
    A.pool_rays = 0;
    A.in_pool = 0;
    A.store_children = 1;
    A.status_tag = 0;
    A.head_out = nullptr;
    const uint64_t nb = (n_in + RPX_BLOCK - 1) / RPX_BLOCK;
    bool timing;
    {
        std::lock_guard<std::mutex> g(g_timing_mutex);
        timing = g_timing_on;
    }
    static thread_local EventSet ev;
    timing = timing && ev.ok;
    const uint32_t variant = scene_variant(scene);
    if (!gauss) {
        if (out->capacity > 0 && (!out->f64 || !out->u32))
            return set_error(RPX_E_INVALID, "output rays have null fields");
        if (timing) HIP_TRY(hipEventRecord(ev.e[0], stream));
        hipLaunchKernelGGL(k_clear, dim3((unsigned)((std::max<uint64_t>(nb, RPX_SHARDS * RPX_LDS_FACES) + 255) / 256)),
                           dim3(256), 0, stream, A.ws.tile_status, nb,
                           A.ws.ctl);
        HIP_TRY(hipGetLastError());
        if (timing) HIP_TRY(hipEventRecord(ev.e[1], stream));
        launch_generation(variant, dim3((unsigned)nb), stream, A);
        HIP_TRY(hipGetLastError());
        if (timing) HIP_TRY(hipEventRecord(ev.e[2], stream));
        GenControl ctl;
        HIP_TRY(finish_generation(ctl, A.ws.ctl, scene->view.faces, scene->view.n_list_faces, face_hits,
                                      stream));
        if (ctl.error) return set_error(RPX_E_HIP, "generation scan: a look-back wait exceeded its poll limit");
        ctl.hits = 0;
        for (int s = 0; s < RPX_SHARDS; ++s) ctl.hits += ctl.hit_shards[s];
        *n_out = ctl.children;
        if (timing) {
            float t_clear = 0, t_gen = 0;
            HIP_TRY(hipEventElapsedTime(&t_clear, ev.e[0], ev.e[1]));
            HIP_TRY(hipEventElapsedTime(&t_gen, ev.e[1], ev.e[2]));
            std::lock_guard<std::mutex> g(g_timing_mutex);
            g_timing_ms[RPX_TIMING_CLEAR] += t_clear;
            g_timing_launches[RPX_TIMING_CLEAR] += 1;
            g_timing_bytes[RPX_TIMING_CLEAR] += (double)nb * 8.0;
            // algorithmic bytes: geometry of every input ray, its length and
            // end face written back, the other 100 bytes of each ray that hit,
            // and every child record
            g_timing_ms[RPX_TIMING_GENERATION] += t_gen;
            g_timing_launches[RPX_TIMING_GENERATION] += 1;
            g_timing_bytes[RPX_TIMING_GENERATION] +=
                (double)n_in * (48.0 + 12.0) + (double)ctl.hits * 100.0 + (double)ctl.children * 188.0;
        }
        if (ctl.children > out->capacity)
            return set_error(RPX_E_CAPACITY, "output capacity " + std::to_string(out->capacity) + " below " +
                                                 std::to_string(ctl.children) + " children");
        return RPX_OK;
    }
    // gausslets: one single-pass launch as well
    if (out->capacity > 0 && (!out->f64 || !out->u32 || !out->para))
        return set_error(RPX_E_INVALID, "output gausslets have null fields");
    if (timing) HIP_TRY(hipEventRecord(ev.e[0], stream));
    hipLaunchKernelGGL(k_clear, dim3((unsigned)((std::max<uint64_t>(nb, RPX_SHARDS * RPX_LDS_FACES) + 255) / 256)),
                       dim3(256), 0, stream, A.ws.tile_status, nb, A.ws.ctl);
    HIP_TRY(hipGetLastError());
    if (timing) HIP_TRY(hipEventRecord(ev.e[1], stream));
    launch_gausslet_generation(variant, dim3((unsigned)nb), stream, A);
    HIP_TRY(hipGetLastError());
    if (timing) HIP_TRY(hipEventRecord(ev.e[2], stream));
    GenControl ctl;
    HIP_TRY(finish_generation(ctl, A.ws.ctl, scene->view.faces, scene->view.n_list_faces, face_hits, stream));
    if (ctl.error) return set_error(RPX_E_HIP, "generation scan: a look-back wait exceeded its poll limit");
    uint64_t hits = 0, parents = 0;
    for (int s2 = 0; s2 < RPX_SHARDS; ++s2) {
        hits += ctl.hit_shards[s2];
        parents += ctl.parent_shards[s2];
    }
    *n_out = ctl.children;
    if (timing) {
        float t_clear = 0, t_gen = 0;
        HIP_TRY(hipEventElapsedTime(&t_clear, ev.e[0], ev.e[1]));
        HIP_TRY(hipEventElapsedTime(&t_gen, ev.e[1], ev.e[2]));
        std::lock_guard<std::mutex> g(g_timing_mutex);
        g_timing_ms[RPX_TIMING_CLEAR] += t_clear;
        g_timing_launches[RPX_TIMING_CLEAR] += 1;
        g_timing_bytes[RPX_TIMING_CLEAR] += (double)nb * 8.0;
        // algorithmic bytes: base and parabasal geometry of every gausslet
        // plus its length and end face, the rest of the base ray (100) and
        // the parabasal normals and lengths (6 x 32) of every gausslet with
        // children, and every child record (188 + 6 x 80)
        g_timing_ms[RPX_TIMING_GAUSSLET] += t_gen;
        g_timing_launches[RPX_TIMING_GAUSSLET] += 1;
        g_timing_bytes[RPX_TIMING_GAUSSLET] += (double)n_in * (48.0 + 6 * 48.0 + 12.0) +
                                               (double)parents * (100.0 + 6 * 32.0) +
                                               (double)ctl.children * (188.0 + 6 * 80.0);
        (void)hits;
    }
    if (ctl.children > out->capacity)
        return set_error(RPX_E_CAPACITY, "output capacity " + std::to_string(out->capacity) + " below " +
                                             std::to_string(ctl.children) + " children");
    return RPX_OK;
}

extern "C" rpx_status_t rpx_trace_segment(const rpx_scene_t *scene, const rpx_rays_t *in, uint64_t n_in,
                                          float max_length, const double *wavelengths,
                                          uint64_t n_wavelengths, rpx_rays_t *out, uint64_t *n_out,
                                          uint32_t *face_hits, void *workspace, uint64_t workspace_bytes,
                                          void *stream) {
    return run_generation(scene, -1, -1, false, in, n_in, (double)max_length, wavelengths, n_wavelengths, out,
                          n_out, face_hits, workspace, workspace_bytes, stream);
}

extern "C" rpx_status_t rpx_trace_one_face_segment(const rpx_scene_t *scene, int32_t face_list,
                                                   int32_t face_idx, const rpx_rays_t *in, uint64_t n_in,
                                                   float max_length, const double *wavelengths,
                                                   uint64_t n_wavelengths, rpx_rays_t *out, uint64_t *n_out,
                                                   uint32_t *face_hits, void *workspace,
                                                   uint64_t workspace_bytes, void *stream) {
    if (!scene || face_list < 0 || face_list >= scene->view.n_face_lists || face_idx < 0)
        return set_error(RPX_E_INVALID, "face list / face index out of range");
    return run_generation(scene, face_list, face_idx, false, in, n_in, (double)max_length, wavelengths,
                          n_wavelengths, out, n_out, face_hits, workspace, workspace_bytes, stream);
}

extern "C" rpx_status_t rpx_trace_gausslet(const rpx_scene_t *scene, int32_t face_list, int32_t face_idx,
                                           const rpx_rays_t *in, uint64_t n_in, double max_length,
                                           const double *wavelengths, uint64_t n_wavelengths, rpx_rays_t *out,
                                           uint64_t *n_out, uint32_t *face_hits, void *workspace,
                                           uint64_t workspace_bytes, void *stream) {
    if (!scene) return set_error(RPX_E_INVALID, "scene is null");
    if (face_list >= scene->view.n_face_lists || (face_list >= 0 && face_idx < 0))
        return set_error(RPX_E_INVALID, "face list / face index out of range");
    if (scene->view.max_children > 2)
        return set_error(RPX_E_UNSUPPORTED, "gausslets through a multi-order grating (its parabasal rays follow "
                         "one order)");
    return run_generation(scene, face_list, face_list >= 0 ? face_idx : -1, true, in, n_in, max_length,
                          wavelengths, n_wavelengths, out, n_out, face_hits, workspace, workspace_bytes, stream);
}

// ------------------------------------------------ device-resident chain
extern "C" uint64_t rpx_chain_workspace_bytes(uint64_t max_rays_per_generation, int32_t max_generations) {
    return chain_workspace_bytes(max_rays_per_generation, max_generations);
}

// rpx_trace_rays: the float64 generations of trace_rays as one device-
// resident chain (run_chain in rpx_gen.h).
extern "C" rpx_status_t rpx_trace_rays(const rpx_scene_t *scene, const rpx_rays_t *in, uint64_t n_in,
                                       float max_length, const double *wavelengths, uint64_t n_wavelengths,
                                       int32_t max_generations, rpx_rays_t *pool, uint64_t *counts,
                                       uint64_t *offsets, int32_t *n_traced, uint32_t *face_hits,
                                       void *workspace, uint64_t workspace_bytes, void *stream_) {
    if (!n_traced || !counts || !offsets) return set_error(RPX_E_INVALID, "counts / offsets / n_traced are null");
    *n_traced = 0;
    counts[0] = n_in;
    offsets[0] = 0;
    if (rpx_status_t st = rpx_check_trace_scene(scene)) return st;
    if (rpx_status_t st = check_rays(in, n_in, false, "input rays")) return st;
    if (!pool || (pool->capacity > 0 && (!pool->f64 || !pool->u32))) return set_error(RPX_E_INVALID, "pool");
    if (max_generations < 0) return set_error(RPX_E_INVALID, "max_generations is negative");
    if (n_in == 0 || max_generations == 0) return RPX_OK;
    if (!wavelengths && n_wavelengths) return set_error(RPX_E_INVALID, "wavelength table is null");
    if (!face_hits && scene->view.n_list_faces > 0) return set_error(RPX_E_INVALID, "face_hits is null");
    ChainSpec spec;
    spec.n_in = n_in;
    spec.pool_rays = pool->capacity;
    spec.max_generations = max_generations;
    spec.max_children = (uint64_t)std::max(scene->view.max_children, 0);
    spec.timing_slot = RPX_TIMING_GENERATION;
    spec.bytes_in = 48.0 + 12.0;    // geometry read, length + end face written
    spec.bytes_hit = 100.0;         // the rest of a ray that hit
    spec.bytes_child = 188.0;       // a child record
    if (rpx_status_t st = check_chain(spec, scene, workspace, workspace_bytes)) return st;
    const uint32_t variant = scene_variant(scene);
    TraceArgs A;
    A.S = scene->view;
    A.out = *pool;
    A.n_in = 0;
    A.max_length = (double)max_length;
    A.wavelengths.wl = wavelengths;
    A.wavelengths.n = wavelengths ? n_wavelengths : 0;
    A.face_hits = nullptr;           // per-face totals come back in the heads
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
        launch_generation(variant, dim3((unsigned)grid), stream, A);
        return hipGetLastError();
    };
    return run_chain(spec, scene, workspace, launch, counts, offsets, n_traced, face_hits, (hipStream_t)stream_);
}

// rpx_trace_gausslets: the gausslet generations of trace_rays as one
// device-resident chain (run_chain in rpx_gen.h).
extern "C" rpx_status_t rpx_trace_gausslets(const rpx_scene_t *scene, const rpx_rays_t *in, uint64_t n_in,
                                       double max_length, const double *wavelengths, uint64_t n_wavelengths,
                                       int32_t max_generations, rpx_rays_t *pool, uint64_t *counts,
                                       uint64_t *offsets, int32_t *n_traced, uint32_t *face_hits,
                                       void *workspace, uint64_t workspace_bytes, void *stream_) {
    if (!n_traced || !counts || !offsets) return set_error(RPX_E_INVALID, "counts / offsets / n_traced are null");
    *n_traced = 0;
    counts[0] = n_in;
    offsets[0] = 0;
    if (rpx_status_t st = rpx_check_trace_scene(scene)) return st;
    if (rpx_status_t st = check_rays(in, n_in, true, "input gausslets")) return st;
    if (!pool || (pool->capacity > 0 && (!pool->f64 || !pool->u32 || !pool->para)))
        return set_error(RPX_E_INVALID, "pool");
    if (max_generations < 0) return set_error(RPX_E_INVALID, "max_generations is negative");
    if (n_in == 0 || max_generations == 0) return RPX_OK;
    if (!wavelengths && n_wavelengths) return set_error(RPX_E_INVALID, "wavelength table is null");
    if (!face_hits && scene->view.n_list_faces > 0) return set_error(RPX_E_INVALID, "face_hits is null");
    ChainSpec spec;
    spec.n_in = n_in;
    spec.pool_rays = pool->capacity;
    spec.max_generations = max_generations;
    spec.max_children = (uint64_t)std::max(scene->view.max_children, 0);
    spec.timing_slot = RPX_TIMING_GAUSSLET;
    // as the one-call-per-generation gausslet path counts them: base and
    // parabasal geometry plus length and end face of every gausslet, the
    // rest of the base ray and the parabasal normals and lengths of every
    // gausslet with children, and every child record
    spec.bytes_in = 48.0 + 6 * 48.0 + 12.0;
    spec.bytes_hit = 100.0 + 6 * 32.0;
    spec.bytes_child = 188.0 + 6 * 80.0;
    spec.bytes_per_parent = true;
    if (rpx_status_t st = check_chain(spec, scene, workspace, workspace_bytes)) return st;
    const uint32_t variant = scene_variant(scene);
    TraceArgs A;
    A.S = scene->view;
    A.out = *pool;
    A.n_in = 0;
    A.max_length = (double)max_length;
    A.wavelengths.wl = wavelengths;
    A.wavelengths.n = wavelengths ? n_wavelengths : 0;
    A.face_hits = nullptr;           // per-face totals come back in the heads
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
        launch_gausslet_generation(variant, dim3((unsigned)grid), stream, A);
        return hipGetLastError();
    };
    return run_chain(spec, scene, workspace, launch, counts, offsets, n_traced, face_hits, (hipStream_t)stream_);
}

// ------------------------------------------------------ AoS <-> SoA
// Packed ray_t: 21 doubles (168 B) then 5 uint32 (20 B) = 188 B = 47 words.
// gausslet_t appends 6 x 10 doubles = 480 B -> 167 words.
// Both directions stage pack_records(WPR) whole records per workgroup in
// LDS: the AoS side moves as one contiguous run of words, and each SoA plane
// as that many consecutive values (doubles as single 8-byte accesses).  The
// record stride (47 or 167 words) is odd, so lanes reading one field of
// consecutive records hit distinct LDS banks.  Gausslet records take 32 per
// workgroup (21 KB of LDS) so that more workgroups fit a CU.
constexpr int pack_records(int wpr) { return wpr > 47 ? 32 : 64; }

template <int WPR, int R = pack_records(WPR)>
__global__ __launch_bounds__(256) void k_pack(rpx_rays_t soa, uint64_t n, uint32_t *aos) {
    __shared__ uint32_t lds[R * WPR];
    const uint64_t r0 = (uint64_t)blockIdx.x * R;
    const int nr = (int)((n - r0) < R ? (n - r0) : R);
    const uint64_t c = soa.capacity;
    for (int k = threadIdx.x; k < 21 * R; k += 256) {
        const int f = k / R, r = k - f * R;
        if (r < nr) {
            const uint64_t v = __double_as_longlong(soa.f64[(uint64_t)f * c + r0 + r]);
            lds[r * WPR + 2 * f] = (uint32_t)v;
            lds[r * WPR + 2 * f + 1] = (uint32_t)(v >> 32);
        }
    }
    for (int k = threadIdx.x; k < 5 * R; k += 256) {
        const int f = k / R, r = k - f * R;
        if (r < nr) lds[r * WPR + 42 + f] = soa.u32[(uint64_t)f * c + r0 + r];
    }
    if constexpr (WPR > 47) {
        for (int k = threadIdx.x; k < 60 * R; k += 256) {
            const int f = k / R, r = k - f * R;
            if (r < nr) {
                const uint64_t v = __double_as_longlong(soa.para[(uint64_t)f * c + r0 + r]);
                lds[r * WPR + 47 + 2 * f] = (uint32_t)v;
                lds[r * WPR + 47 + 2 * f + 1] = (uint32_t)(v >> 32);
            }
        }
    }
    __syncthreads();
    uint32_t *dst = aos + r0 * WPR;
    for (int k = threadIdx.x; k < nr * WPR; k += 256) dst[k] = lds[k];
}

template <int WPR, int R = pack_records(WPR)>
__global__ __launch_bounds__(256) void k_unpack(const uint32_t *aos, uint64_t n, rpx_rays_t soa) {
    __shared__ uint32_t lds[R * WPR];
    const uint64_t r0 = (uint64_t)blockIdx.x * R;
    const int nr = (int)((n - r0) < R ? (n - r0) : R);
    const uint32_t *src = aos + r0 * WPR;
    for (int k = threadIdx.x; k < nr * WPR; k += 256) lds[k] = src[k];
    __syncthreads();
    const uint64_t c = soa.capacity;
    for (int k = threadIdx.x; k < 21 * R; k += 256) {
        const int f = k / R, r = k - f * R;
        if (r < nr) {
            const uint64_t v = (uint64_t)lds[r * WPR + 2 * f] | ((uint64_t)lds[r * WPR + 2 * f + 1] << 32);
            soa.f64[(uint64_t)f * c + r0 + r] = __longlong_as_double((long long)v);
        }
    }
    for (int k = threadIdx.x; k < 5 * R; k += 256) {
        const int f = k / R, r = k - f * R;
        if (r < nr) soa.u32[(uint64_t)f * c + r0 + r] = lds[r * WPR + 42 + f];
    }
    if constexpr (WPR > 47) {
        for (int k = threadIdx.x; k < 60 * R; k += 256) {
            const int f = k / R, r = k - f * R;
            if (r < nr) {
                const uint64_t v =
                    (uint64_t)lds[r * WPR + 47 + 2 * f] | ((uint64_t)lds[r * WPR + 47 + 2 * f + 1] << 32);
                soa.para[(uint64_t)f * c + r0 + r] = __longlong_as_double((long long)v);
            }
        }
    }
}

extern "C" rpx_status_t rpx_pack_rays(const rpx_rays_t *soa, uint64_t n, void *aos, int32_t gausslet,
                                      void *stream) {
    rpx_status_t st = check_rays(soa, n, gausslet != 0, "rays");
    if (st) return st;
    if (n == 0) return RPX_OK;
    if (!aos) return set_error(RPX_E_INVALID, "AoS buffer is null");
    const uint64_t recs = (uint64_t)pack_records(gausslet ? 167 : 47);
    const uint64_t nb = (n + recs - 1) / recs;
    if (nb > 0x7fffffffull) return set_error(RPX_E_INVALID, "too many rays for one pack launch");
    if (gausslet)
        hipLaunchKernelGGL(k_pack<167>, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream, *soa, n,
                           (uint32_t *)aos);
    else
        hipLaunchKernelGGL(k_pack<47>, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream, *soa, n,
                           (uint32_t *)aos);
    HIP_TRY(hipGetLastError());
    return RPX_OK;
}

extern "C" rpx_status_t rpx_unpack_rays(const void *aos, uint64_t n, rpx_rays_t *soa, int32_t gausslet,
                                        void *stream) {
    rpx_status_t st = check_rays(soa, n, gausslet != 0, "rays");
    if (st) return st;
    if (n == 0) return RPX_OK;
    if (!aos) return set_error(RPX_E_INVALID, "AoS buffer is null");
    const uint64_t recs = (uint64_t)pack_records(gausslet ? 167 : 47);
    const uint64_t nb = (n + recs - 1) / recs;
    if (nb > 0x7fffffffull) return set_error(RPX_E_INVALID, "too many rays for one unpack launch");
    if (gausslet)
        hipLaunchKernelGGL(k_unpack<167>, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream,
                           (const uint32_t *)aos, n, *soa);
    else
        hipLaunchKernelGGL(k_unpack<47>, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream,
                           (const uint32_t *)aos, n, *soa);
    HIP_TRY(hipGetLastError());
    return RPX_OK;
}
