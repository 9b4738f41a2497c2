// Host-side handle types behind the C ABI's opaque pointers.
#pragma once

#include <hip/hip_runtime.h>

#include "lr_internal.h"
#include "trace_kernels.h"

struct lr_device {
    int ordinal = 0;
    hipStream_t stream = nullptr;
    int compute_units = 0;
    // This device's copy of the allocation-free entry point's scratch array
    // (its device address; null, with scratch_error set, when the runtime
    // could not give it).
    void *scratch = nullptr;
    int scratch_error = 0;
};

namespace lr {
// Looks up the scratch array's address on `device` (current), once, at open.
void scratch_attach(lr_device *device);
}  // namespace lr

struct lr_gpu_scene {
    lr_device *device = nullptr;
    // Layouts: widths 2, 4, 8 with float planes, then widths 4 and 8
    // quantized to bytes (traversal layout codes 2, 4, 8, 20, 24).
    void *wide[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    void *tris = nullptr;                         // 48 B per leaf slot
    void *faces = nullptr;
    void *shading = nullptr;
    void *uvs = nullptr;
    void *spill = nullptr;
    uint64_t spill_entries = 0;
    uint32_t *work = nullptr;  // the persistent launches' two counters
    void *ranges = nullptr;    // width-2 nodes' triangle slot ranges (uint2 each)
    // Kept for refit: the binary tree, leaf-order point indices, the point
    // count; made at the first refit: parents, arrival counters and each wide
    // layout's binary node per child slot.
    lr_bvh_node *binary = nullptr;
    uint32_t *indices = nullptr;
    size_t points = 0;
    uint32_t *parent = nullptr, *arrive = nullptr;
    uint32_t *slots[3] = {nullptr, nullptr, nullptr};
    float root_lo[3] = {0, 0, 0};
    float root_hi[3] = {0, 0, 0};
    // Gaussian splats: records (16 words each, leaf order), harmonics words
    // and their binary tree (32-byte nodes), as uploaded.
    void *splat_records = nullptr;
    void *splat_words = nullptr;
    void *splat_nodes = nullptr;
    size_t splat_count = 0;
    lr_scene_info info{};
    lr_trace_config config{20, 64, 16, 32};

    ~lr_gpu_scene();
    int ensure_spill(uint64_t threads, uint32_t stack);
    int config_width_index() const;
    lr::DeviceScene device_scene() const;
    const lr::TraceVariant &variant() const;
};

namespace lr {
using GpuScene = lr_gpu_scene;
}

namespace lr {
// Device BVH build (bvh_build_gpu.hip): `n` triangles whose points and
// indices are in device memory; `out` (room for max(1, 2n) nodes) and `order`
// in device memory receive the reference's tree and leaf order.
int build_bvh_device(lr_device *dev, const float *positions, const uint32_t *indices, size_t n, lr_bvh_node *out,
                     uint32_t *order, size_t *node_count);
// Device collapse (bvh_collapse_gpu.hip) of the binary tree in device memory.
int collapse_wide_device(lr_device *dev, const lr_bvh_node *nodes, size_t node_count, bool empty, int width,
                         DeviceLayout &layout, bool with_slots = false);
// The quantized layout (64 B a node at width 4, 128 B at 8) of a float
// layout of `count` nodes in device memory; allocates *dst when null.
int quantize_wide_device(lr_device *dev, const void *src, size_t count, int width, void **dst);
// Layout index of a traversal layout code (2, 4, 8, 20, 24), and the index of
// its stack bound (the width's), or -1.
inline int layout_index(uint32_t code) {
    return code == 2 ? 0 : code == 4 ? 1 : code == 8 ? 2 : code == 20 ? 3 : code == 24 ? 4 : -1;
}
inline int bound_index(uint32_t code) {
    int l = layout_index(code);
    return l < 0 ? -1 : (l < 3 ? l : l - 2);
}
// Refit (bvh_refit_gpu.hip): new points in device memory, same triangles.
int refit_scene_device(lr_gpu_scene *scene, const float *positions);
// Leaf-order triangles (p0, e1, e2 as 3 float4) from device points and indices.
void pack_triangles_device(lr_device *dev, const float *positions, const uint32_t *indices, size_t n, float *tris);
// dst[t * words + w] = src[order[t] * words + w]: per-triangle rows into leaf order.
void gather_rows_device(lr_device *dev, const uint32_t *src, const uint32_t *order, size_t n, uint32_t words,
                        uint32_t *dst);
}  // namespace lr
