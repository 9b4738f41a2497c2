// Refit: the scene's points move, its triangles and tree stay. Every binary
// node's box is recomputed bottom-up (a leaf grows an empty box by its
// triangles' corners, as the build does; an interior node takes the union of
// its children's boxes: the same values min/max give in any order), each
// wide node keeps the children it was built with and takes their new boxes,
// and the leaf-order triangles are repacked. No reference counterpart: the
// reference rebuilds the scene on any change.
//
// Bottom-up without levels: a thread per leaf walks towards the root; at each
// interior node the first of the two children to arrive stops, the second
// (which then sees both children's boxes) writes the node and goes on.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lr_internal.h"
#include "lr_scene.h"

namespace lr {
namespace grefit {

constexpr uint32_t kNone = 0xffffffffu;

__device__ __forceinline__ float mn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float mx(float a, float b) { return a > b ? a : b; }

__global__ __launch_bounds__(256) void parent_kernel(const lr_bvh_node *nodes, uint32_t count, uint32_t *parent) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= count) return;
    if (i == 0) parent[0] = kNone;
    const lr_bvh_node n = nodes[i];
    if (n.count == 0 && n.first + 1u < count) {
        parent[n.first] = i;
        parent[n.first + 1] = i;
    }
}

__global__ __launch_bounds__(256) void refit_kernel(lr_bvh_node *nodes, uint32_t count, const uint32_t *parent,
                                                   uint32_t *arrive, const float *positions, const uint32_t *indices) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= count) return;
    lr_bvh_node n = nodes[i];
    if (n.count == 0) return;
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (uint32_t t = n.first; t < n.first + n.count; t++) {
        float tlo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, thi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
        for (int k = 0; k < 3; k++) {
            const float *p = positions + (size_t)indices[(size_t)t * 3 + k] * 3;
            for (int d = 0; d < 3; d++) {
                tlo[d] = mn(tlo[d], p[d]);
                thi[d] = mx(thi[d], p[d]);
            }
        }
        for (int d = 0; d < 3; d++) {
            lo[d] = mn(lo[d], tlo[d]);
            hi[d] = mx(hi[d], thi[d]);
        }
    }
    for (int d = 0; d < 3; d++) nodes[i].lo[d] = lo[d], nodes[i].hi[d] = hi[d];
    // Up the tree: the second child to arrive finishes the parent.
    uint32_t at = parent[i];
    while (at != kNone) {
        __threadfence();
        if (atomicAdd(&arrive[at], 1u) == 0u) return;
        __threadfence();
        const uint32_t c = nodes[at].first;
        // Loads through the device-coherent level: the sibling was written by
        // another thread, possibly on another compute unit.
        float a[6], b[6];
        for (int d = 0; d < 3; d++) {
            a[d] = __hip_atomic_load(&nodes[c].lo[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            a[3 + d] = __hip_atomic_load(&nodes[c].hi[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            b[d] = __hip_atomic_load(&nodes[c + 1].lo[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            b[3 + d] = __hip_atomic_load(&nodes[c + 1].hi[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        for (int d = 0; d < 3; d++) {
            nodes[at].lo[d] = mn(a[d], b[d]);
            nodes[at].hi[d] = mx(a[3 + d], b[3 + d]);
        }
        at = parent[at];
    }
}

// Child slot boxes of width-W nodes from the binary nodes they hold.
template <int W>
__global__ __launch_bounds__(256) void slots_kernel(const lr_bvh_node *nodes, const uint32_t *slots, uint32_t wide,
                                                   uint8_t *out, uint32_t node_bytes) {
    uint64_t k = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (k >= (uint64_t)wide * W) return;
    uint32_t kid = slots[k];
    if (kid == kNone) return;
    const uint32_t w = (uint32_t)(k / W), i = (uint32_t)(k % W);
    WideNode<W> *node = reinterpret_cast<WideNode<W> *>(out + (size_t)w * node_bytes);
    const lr_bvh_node n = nodes[kid];
    node->lo_x[i] = n.lo[0];
    node->lo_y[i] = n.lo[1];
    node->lo_z[i] = n.lo[2];
    node->hi_x[i] = n.hi[0];
    node->hi_y[i] = n.hi[1];
    node->hi_z[i] = n.hi[2];
}

}  // namespace grefit

int refit_scene_device(lr_gpu_scene *s, const float *positions) {
    lr_device *dev = s->device;
    hipStream_t st = dev->stream;
    hipError_t e;
    const uint32_t count = (uint32_t)s->info.binary_nodes;
    const dim3 grid((count + 255) / 256);
    // First refit: parents, counters and the slot maps of the layouts as
    // built (the tree still holds the boxes they were built from).
    if (!s->parent) {
        uint32_t *slots[3] = {nullptr, nullptr, nullptr}, *parent = nullptr, *arrive = nullptr;
        auto drop = [&] {
            for (void *p : {(void *)slots[0], (void *)slots[1], (void *)slots[2], (void *)parent, (void *)arrive})
                if (p) (void)hipFree(p);
        };
        const int widths[3] = {2, 4, 8};
        for (int wi = 0; wi < 3; wi++) {
            DeviceLayout layout;
            int rc = collapse_wide_device(dev, s->binary, count, false, widths[wi], layout, true);
            if (rc) {
                drop();
                return rc;
            }
            const bool same = layout.nodes == s->info.wide_nodes[wi];
            slots[wi] = static_cast<uint32_t *>(layout.slots);
            for (void *p : {layout.data, layout.ranges})
                if (p) (void)hipFree(p);
            if (!same) {
                drop();
                return fail(LR_DEVICE_ERROR, "refit: the collapse no longer matches the uploaded layout");
            }
        }
        if ((e = hipMalloc((void **)&parent, (size_t)count * 4)) != hipSuccess ||
            (e = hipMalloc((void **)&arrive, (size_t)count * 4)) != hipSuccess) {
            drop();
            return hip_fail(e, "hipMalloc(refit)");
        }
        hipLaunchKernelGGL(grefit::parent_kernel, grid, dim3(256), 0, st, s->binary, count, parent);
        for (int wi = 0; wi < 3; wi++) s->slots[wi] = slots[wi];
        s->arrive = arrive;
        s->parent = parent;  // set last: the refit state is complete
    }
    if ((e = hipMemsetAsync(s->arrive, 0, (size_t)count * 4, st)) != hipSuccess) return hip_fail(e, "hipMemset(refit)");
    hipLaunchKernelGGL(grefit::refit_kernel, grid, dim3(256), 0, st, s->binary, count, s->parent, s->arrive, positions,
                       s->indices);
    for (int wi = 0; wi < 3; wi++) {
        const uint32_t w = (uint32_t)s->info.wide_nodes[wi];
        const uint32_t node_bytes = (uint32_t)(s->info.wide_bytes[wi] / std::max<size_t>(1, w));
        const uint64_t items = (uint64_t)w * (2u << wi);
        const dim3 g((uint32_t)((items + 255) / 256));
        uint8_t *out = static_cast<uint8_t *>(s->wide[wi]);
        if (wi == 0) hipLaunchKernelGGL(grefit::slots_kernel<2>, g, dim3(256), 0, st, s->binary, s->slots[0], w, out, node_bytes);
        if (wi == 1) hipLaunchKernelGGL(grefit::slots_kernel<4>, g, dim3(256), 0, st, s->binary, s->slots[1], w, out, node_bytes);
        if (wi == 2) hipLaunchKernelGGL(grefit::slots_kernel<8>, g, dim3(256), 0, st, s->binary, s->slots[2], w, out, node_bytes);
    }
    // The quantized layouts follow their float layouts.
    int rc;
    if ((rc = quantize_wide_device(dev, s->wide[1], s->info.wide_nodes[1], 4, &s->wide[3])) ||
        (rc = quantize_wide_device(dev, s->wide[2], s->info.wide_nodes[2], 8, &s->wide[4])))
        return rc;
    pack_triangles_device(dev, positions, s->indices, s->info.triangles, static_cast<float *>(s->tris));
    lr_bvh_node root;
    if ((e = hipMemcpyAsync(&root, s->binary, sizeof(root), hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess)
        return hip_fail(e, "refit");
    for (int d = 0; d < 3; d++) s->root_lo[d] = root.lo[d], s->root_hi[d] = root.hi[d];
    return 0;
}

}  // namespace lr
