// Capture-face selection: select_ray_intersections /
// select_gausslet_intersections (ctracer.pyx:1981-2058) for one collection.
//
// Every ray (the base ray of every gausslet) is tested along its own segment
// origin -> origin + direction * length against one face list, exactly as
// FaceList.intersect_c (ctracer.pyx:1882-1904) does it; a ray whose nearest
// accepted face has a global index >= 0 is copied to the output with the new
// length and end face and with wavelength_idx mapped through the caller's
// table.  The input is not modified.  Copies keep input order: one launch
// places them with the decoupled look-back over tiles of RPX_BLOCK rays used
// by the generation kernels.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <string>
#include "rpx_materials.h"
#include "rpx_host.h"
#include "rpx_gen.h"

using namespace rpx;

struct SelectArgs {
    SceneView S;
    int32_t list;
    rpx_rays_t in;
    rpx_rays_t out;
    uint64_t n_in;
    uint64_t out_offset;
    const uint32_t *wl_map;
    uint64_t n_wl_map;
    Workspace ws;   // ctl->hits counts selected rays whose wavelength_idx is outside wl_map
};

template <bool GAUSS, uint32_t F>
__global__ __launch_bounds__(RPX_BLOCK) void k_select(SelectArgs A) {
    __shared__ SceneLds lds_scene;
    __shared__ uint32_t lds_wave[RPX_BLOCK / 64];
    __shared__ uint32_t lds_tile;
    __shared__ uint64_t lds_excl;
    const SceneView S = stage_scene(A.S, lds_scene);
    if (threadIdx.x == 0) lds_tile = atomicAdd(&A.ws.ctl->tile_counter, 1u);
    __syncthreads();
    const uint32_t tile = lds_tile;
    const uint64_t i = (uint64_t)tile * RPX_BLOCK + threadIdx.x;
    bool sel = false;
    double length = 0.0;
    uint32_t end_face = RPX_NO_FACE;
    if (i < A.n_in) {
        const double *f = A.in.f64;
        const uint64_t c = A.in.capacity;
        const V3 origin = v3(f[RPX_OX * c + i], f[RPX_OY * c + i], f[RPX_OZ * c + i]);
        const V3 dir = v3(f[RPX_DX * c + i], f[RPX_DY * c + i], f[RPX_DZ * c + i]);
        length = f[RPX_LENGTH * c + i];
        const V3 point = add(origin, mul(dir, length));
        const rpx_face_list_t &fl = S.face_lists[A.list];
        const V3 p1 = transform(fl.inv_trans, origin);
        const V3 p2 = transform(fl.inv_trans, point);
        for (int k = fl.first_face; k < fl.first_face + fl.n_faces; ++k) {
            const rpx_face_t &fc = S.faces[k];
            const double dist = face_intersect<F>(S, fc, p1, p2, 1);
            if (fc.tolerance < dist && dist < length) {
                length = dist;
                end_face = (uint32_t)fc.global_idx;
                sel = fc.global_idx >= 0;
            }
        }
    }
    uint32_t block_total;
    const uint32_t local = block_scan_small(sel ? 1u : 0u, lds_wave, &block_total);
    if (threadIdx.x < 64) {
        const uint64_t excl = lookback(A.ws.tile_status, tile, 0, block_total, &A.ws.ctl->error);
        if (threadIdx.x == 0) lds_excl = excl;
    }
    __syncthreads();
    if (sel) {
        const uint64_t pos = A.out_offset + lds_excl + local;
        Ray r = load_ray(A.in, i);
        r.length = length;
        r.end_face_idx = end_face;
        if (r.wl_idx < A.n_wl_map) {
            r.wl_idx = A.wl_map[r.wl_idx];
        } else {
            atomicAdd((unsigned long long *)&A.ws.ctl->hits, 1ull);
        }
        if (pos < A.out.capacity) {
            store_ray(A.out, pos, r);
            if (GAUSS) {
                const uint64_t ci = A.in.capacity, co = A.out.capacity;
                for (int q = 0; q < RPX_N_PARA * RPX_PARA_FIELDS; ++q)
                    A.out.para[(uint64_t)q * co + pos] = A.in.para[(uint64_t)q * ci + i];
            }
        }
    }
    if (threadIdx.x == 0 && (uint64_t)tile == (A.n_in - 1) / RPX_BLOCK) A.ws.ctl->children = lds_excl + block_total;
}

template <uint32_t F>
static void launch_select_f(bool gauss, dim3 grid, hipStream_t stream, const SelectArgs &A) {
    if (gauss)
        hipLaunchKernelGGL((k_select<true, F>), grid, dim3(RPX_BLOCK), 0, stream, A);
    else
        hipLaunchKernelGGL((k_select<false, F>), grid, dim3(RPX_BLOCK), 0, stream, A);
}

extern "C" rpx_status_t rpx_select_intersections(const rpx_scene_t *scene, int32_t face_list,
                                                 const rpx_rays_t *in, uint64_t n_in, int32_t gausslets,
                                                 const uint32_t *wl_map, uint64_t n_wl_map, rpx_rays_t *out,
                                                 uint64_t out_offset, uint64_t *n_out, void *workspace,
                                                 uint64_t workspace_bytes, void *stream_) {
    if (!scene) return rpx_set_error(RPX_E_INVALID, "scene is null");
    if (!n_out) return rpx_set_error(RPX_E_INVALID, "n_out is null");
    *n_out = 0;
    if (face_list < 0 || face_list >= scene->view.n_face_lists)
        return rpx_set_error(RPX_E_INVALID, "face list out of range");
    const bool gauss = gausslets != 0;
    if (!in) return rpx_set_error(RPX_E_INVALID, "input rays are null");
    if (in->capacity < n_in) return rpx_set_error(RPX_E_INVALID, "input rays: capacity below count");
    if (n_in > 0 && (!in->f64 || !in->u32 || (gauss && !in->para)))
        return rpx_set_error(RPX_E_INVALID, "input rays have null fields");
    if (!out) return rpx_set_error(RPX_E_INVALID, "output rays are null");
    if (out->capacity > 0 && (!out->f64 || !out->u32 || (gauss && !out->para)))
        return rpx_set_error(RPX_E_INVALID, "output rays have null fields");
    if (n_wl_map > 0 && !wl_map) return rpx_set_error(RPX_E_INVALID, "wavelength map is null");
    if (n_in == 0) return RPX_OK;
    if (n_in > 0xFFFFFFFFull) return rpx_set_error(RPX_E_UNSUPPORTED, "more than 2^32-1 rays in one collection");
    if (!workspace || workspace_bytes < rpx_trace_workspace_bytes(n_in))
        return rpx_set_error(RPX_E_INVALID, "workspace too small");
    hipStream_t stream = (hipStream_t)stream_;
    SelectArgs A;
    A.S = scene->view;
    A.list = face_list;
    A.in = *in;
    A.out = *out;
    A.n_in = n_in;
    A.out_offset = out_offset;
    A.wl_map = wl_map;
    A.n_wl_map = n_wl_map;
    A.ws = carve(workspace, n_in);
    const uint64_t nb = (n_in + RPX_BLOCK - 1) / RPX_BLOCK;
    hipLaunchKernelGGL(k_clear, dim3((unsigned)((std::max<uint64_t>(nb, RPX_SHARDS * RPX_LDS_FACES) + 255) / 256)),
                       dim3(256), 0, stream, A.ws.tile_status, nb, A.ws.ctl);
    RPX_HIP_TRY(hipGetLastError());
    const uint32_t variant = scene_variant(scene);
    if (variant == V_SPHERES) launch_select_f<V_SPHERES>(gauss, dim3((unsigned)nb), stream, A);
    else if (variant == V_LENS) launch_select_f<V_LENS>(gauss, dim3((unsigned)nb), stream, A);
    else launch_select_f<V_ALL>(gauss, dim3((unsigned)nb), stream, A);
    RPX_HIP_TRY(hipGetLastError());
    GenControl ctl;
    RPX_HIP_TRY(read_control(ctl, A.ws.ctl, stream));
    if (ctl.error) return rpx_set_error(RPX_E_HIP, "selection scan: a look-back wait exceeded its poll limit");
    *n_out = ctl.children;
    if (ctl.hits)
        return rpx_set_error(RPX_E_INVALID, std::to_string(ctl.hits) + " selected rays have a wavelength_idx outside "
                                        "the wavelength map");
    if (out_offset + ctl.children > out->capacity)
        return rpx_set_error(RPX_E_CAPACITY, "output capacity " + std::to_string(out->capacity) + " below " +
                                             std::to_string(out_offset + ctl.children) + " rays");
    return RPX_OK;
}

// ------------------------------------------------------------- capture
// ResampleGaussletMaterial.eval_child_ray_c (cmaterials.pyx:1811-1822) appends
// every gausslet whose base ray ends on one of its faces to captured_rays,
// after FaceList.intersect_c set the base ray's length and end face but before
// trace_parabasal_rays (ctracer.pyx:2263-2268) moves the parabasal lengths.
// Run after the generation: the base ray comes from `in`, the six parabasal
// lengths from `para_length` (row j at para_length[j * stride + i]), taken
// before the generation.  Output in input order at out[out_offset + k].
#define RPX_CAPTURE_MAX_FACES 16
struct CaptureArgs {
    rpx_rays_t in;
    rpx_rays_t out;
    uint64_t n_in;
    uint64_t out_offset;
    const double *para_length;
    uint64_t stride;
    uint32_t faces[RPX_CAPTURE_MAX_FACES];
    int32_t n_faces;
    Workspace ws;
};

__global__ __launch_bounds__(RPX_BLOCK) void k_capture(CaptureArgs A) {
    __shared__ uint32_t lds_wave[RPX_BLOCK / 64];
    __shared__ uint32_t lds_tile;
    __shared__ uint64_t lds_excl;
    if (threadIdx.x == 0) lds_tile = atomicAdd(&A.ws.ctl->tile_counter, 1u);
    __syncthreads();
    const uint32_t tile = lds_tile;
    const uint64_t i = (uint64_t)tile * RPX_BLOCK + threadIdx.x;
    bool sel = false;
    if (i < A.n_in) {
        const uint32_t ef = A.in.u32[RPX_END_FACE_IDX * A.in.capacity + i];
        for (int k = 0; k < A.n_faces; ++k) sel = sel || ef == A.faces[k];
    }
    uint32_t block_total;
    const uint32_t local = block_scan_small(sel ? 1u : 0u, lds_wave, &block_total);
    if (threadIdx.x < 64) {
        const uint64_t excl = lookback(A.ws.tile_status, tile, 0, block_total, &A.ws.ctl->error);
        if (threadIdx.x == 0) lds_excl = excl;
    }
    __syncthreads();
    const uint64_t pos = A.out_offset + lds_excl + local;
    if (sel && pos < A.out.capacity) {
        store_ray(A.out, pos, load_ray(A.in, i));
        const uint64_t ci = A.in.capacity, co = A.out.capacity;
        for (int j = 0; j < RPX_N_PARA; ++j) {
            const uint64_t b = (uint64_t)j * RPX_PARA_FIELDS;
            for (int q = 0; q < 9; ++q) A.out.para[(b + q) * co + pos] = A.in.para[(b + q) * ci + i];
            A.out.para[(b + 9) * co + pos] =
                A.para_length ? A.para_length[(uint64_t)j * A.stride + i] : A.in.para[(b + 9) * ci + i];
        }
    }
    if (threadIdx.x == 0 && (uint64_t)tile == (A.n_in - 1) / RPX_BLOCK) A.ws.ctl->children = lds_excl + block_total;
}

extern "C" rpx_status_t rpx_capture_gausslets(const rpx_rays_t *in, uint64_t n_in, const int32_t *face_ids,
                                              int32_t n_face_ids, const double *para_length,
                                              uint64_t para_length_stride, rpx_rays_t *out, uint64_t out_offset,
                                              uint64_t *n_out, void *workspace, uint64_t workspace_bytes,
                                              void *stream_) {
    if (!n_out) return rpx_set_error(RPX_E_INVALID, "n_out is null");
    *n_out = 0;
    if (!in || in->capacity < n_in) return rpx_set_error(RPX_E_INVALID, "input gausslets");
    if (n_in > 0 && (!in->f64 || !in->u32 || !in->para)) return rpx_set_error(RPX_E_INVALID, "input gausslets");
    if (!out || (out->capacity > 0 && (!out->f64 || !out->u32 || !out->para)))
        return rpx_set_error(RPX_E_INVALID, "output gausslets");
    if (n_face_ids < 0 || n_face_ids > RPX_CAPTURE_MAX_FACES || (n_face_ids > 0 && !face_ids))
        return rpx_set_error(RPX_E_INVALID, "face ids (at most 16)");
    if (para_length && para_length_stride < n_in) return rpx_set_error(RPX_E_INVALID, "para length stride");
    if (n_in == 0 || n_face_ids == 0) return RPX_OK;
    if (n_in > 0xFFFFFFFFull) return rpx_set_error(RPX_E_UNSUPPORTED, "more than 2^32-1 gausslets");
    if (!workspace || workspace_bytes < rpx_trace_workspace_bytes(n_in))
        return rpx_set_error(RPX_E_INVALID, "workspace too small");
    hipStream_t stream = (hipStream_t)stream_;
    CaptureArgs A;
    A.in = *in;
    A.out = *out;
    A.n_in = n_in;
    A.out_offset = out_offset;
    A.para_length = para_length;
    A.stride = para_length_stride;
    A.n_faces = n_face_ids;
    for (int k = 0; k < RPX_CAPTURE_MAX_FACES; ++k) A.faces[k] = k < n_face_ids ? (uint32_t)face_ids[k] : 0u;
    A.ws = carve(workspace, n_in);
    const uint64_t nb = (n_in + RPX_BLOCK - 1) / RPX_BLOCK;
    hipLaunchKernelGGL(k_clear, dim3((unsigned)((std::max<uint64_t>(nb, RPX_SHARDS * RPX_LDS_FACES) + 255) / 256)),
                       dim3(256), 0, stream, A.ws.tile_status, nb, A.ws.ctl);
    RPX_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_capture, dim3((unsigned)nb), dim3(RPX_BLOCK), 0, stream, A);
    RPX_HIP_TRY(hipGetLastError());
    GenControl ctl;
    RPX_HIP_TRY(read_control(ctl, A.ws.ctl, stream));
    if (ctl.error) return rpx_set_error(RPX_E_HIP, "capture scan: a look-back wait exceeded its poll limit");
    *n_out = ctl.children;
    if (out_offset + ctl.children > out->capacity)
        return rpx_set_error(RPX_E_CAPACITY, "output capacity " + std::to_string(out->capacity) + " below " +
                                                 std::to_string(out_offset + ctl.children) + " gausslets");
    return RPX_OK;
}
