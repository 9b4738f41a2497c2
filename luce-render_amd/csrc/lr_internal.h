// Shared host-side declarations of the luce-render HIP library.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/luce_render_hip.h"

namespace lr {

// Sets the calling thread's error message and returns `code`.
int fail(int code, const std::string &message);
// Maps a HIP status to LR_DEVICE_ERROR with the call named.
int hip_fail(int status, const char *what);

// ---------------------------------------------------------------- BVH build

// Binned-SAH binary BVH with the reference's split rules (src/render/bvh.lucb:
// 29-182), built by subtree tasks on `threads` host threads, then numbered in
// the reference's pair-allocation order. Returns the node count.
size_t build_binary_bvh(const float *positions, const uint32_t *indices, size_t triangles, int threads,
                        lr_bvh_node *nodes, uint32_t *order);
// The same build over primitive boxes (six floats each: low, high), as
// build_tree (bvh.lucb:85) takes them for the splats' tree.
size_t build_box_bvh(const float *boxes, size_t count, int threads, lr_bvh_node *nodes, uint32_t *order);

// --------------------------------------------------------------- wide BVH

// A node of width W, 64-byte aligned: child boxes as lo/hi planes per axis,
// then child references. A child reference is a node index (interior) or a
// leaf-order triangle slot with `count` > 0 (leaf); count 0 and ref
// 0xffffffff mark an empty child.
template <int W> struct WideNode {
    float lo_x[W], hi_x[W], lo_y[W], hi_y[W], lo_z[W], hi_z[W];
    uint32_t child[W];
    uint8_t count[W];
};

struct WideLayout {
    int width = 0;
    size_t node_bytes = 0;  // per node, padded to a multiple of 64
    size_t nodes = 0;
    std::vector<uint8_t> bytes;
    uint32_t stack_bound = 0;
    // Width 2 (the binary tree itself): each node's range of leaf-order
    // triangle slots, lowest and highest, for the wider traversals' ties.
    std::vector<uint32_t> ranges;
};

// Collapses the binary tree into width-W nodes (W = 2, 4, 8): each wide node
// holds the binary descendants reached by opening the largest-area interior
// child until W are held. Boxes are the binary nodes' own, so a child's box
// test is the one the binary traversal makes for that node.
WideLayout collapse_wide(const lr_bvh_node *nodes, size_t node_count, int width);

// ------------------------------------------------------ device-side repacking

// A wide layout in device memory (collapse_wide's bytes, built on the GPU).
struct DeviceLayout {
    void *data = nullptr;    // nodes * node_bytes
    void *ranges = nullptr;  // width 2: uint2 slot range per node
    void *slots = nullptr;   // when asked: the binary node in each child slot (width per node)
    size_t node_bytes = 0, nodes = 0, bytes = 0;
    uint32_t stack_bound = 0;
};

}  // namespace lr
