// Host-side internals shared by the library's translation units.
#pragma once
#include <string>
#include <vector>
#include "rpx_internal.h"

struct rpx_scene {
    rpx::SceneView view;
    void *device_block;       // the device tables, a block of the device's scene heap
    int device;               // the device they live on
    int32_t n_global_faces;
    int32_t list_sizes[16];
    uint32_t features;        // face / material families present (F_*, M_*)
    bool negative_face_idx;   // a face has global_idx < 0: usable for selection, not for tracing
    std::vector<rpx_shape_op_t> host_shape_ops;   // for validating query roots
    std::vector<int32_t> face_global_idx;         // Face.idx of each face record (host copy)
};

// Device memory for an rpx::BlockHeap (rpx_heap.h): chunks from hipMalloc,
// returned with hipFree (which waits for the device, so heaps return chunks
// rarely).
struct DeviceChunks {
    void *get(uint64_t bytes) {
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        return p;
    }
    void put(void *p) { (void)hipFree(p); }
};

// Checks that ops[begin..root] is a well-formed postfix program.
rpx_status_t rpx_check_shape_program(const rpx_shape_op_t *ops, int n_ops, int root);

// Kernel timing totals (rpx_timing_enable / rpx_timing_read): whether timing
// is on, and adding one launch of `ms` milliseconds and `bytes` algorithmic
// bytes to slot `slot` (RPX_TIMING_*).
bool rpx_timing_is_on();
void rpx_timing_add(int slot, float ms, double bytes);

// Records the message for rpx_last_error() and returns `code`.
rpx_status_t rpx_set_error(rpx_status_t code, const std::string &msg);

// Launch-order tiles of the device-resident generation chains: the bound on
// one look-back wait in real-time-counter ticks (0: the default bound; set by
// rpx_debug_lookback_ticks), and the count of chains that ran again with
// ticket order after a wait outlasted it.
uint64_t rpx_lookback_ticks_override();
void rpx_note_tile_order_fallback();

// RPX_OK if the scene can be traced (every face has a global index >= 0,
// which indexes the per-face hit counters), else RPX_E_INVALID.
rpx_status_t rpx_check_trace_scene(const rpx_scene_t *scene);

#define RPX_HIP_TRY(expr)                                                                     \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            return rpx_set_error(RPX_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)
