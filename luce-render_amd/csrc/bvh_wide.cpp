// Repacks the reference's binary BVH (32-byte nodes, bvh.lucb:23-27) into the
// traversal layouts the kernels read: width-W nodes that carry their children's
// boxes, padded to whole 64-byte lines (64 B for W=2, 128 B for W=4, 256 B for
// W=8), so one node is one or two 128-byte HBM requests per lane.
#include <algorithm>
#include <cstring>

#include "lr_internal.h"

namespace lr {
namespace {

float node_area(const lr_bvh_node &n) {
    float x = n.hi[0] - n.lo[0], y = n.hi[1] - n.lo[1], z = n.hi[2] - n.lo[2];
    if (x < 0.0f || y < 0.0f || z < 0.0f) return 0.0f;
    return 2.0f * (x * y + y * z + z * x);
}

template <int W> size_t padded_bytes() { return (sizeof(WideNode<W>) + 63) / 64 * 64; }

// The binary descendants a wide node over binary node `b` holds: b's two
// children, then the largest-area interior one opened in place until W.
int gather(const lr_bvh_node *nodes, uint32_t b, int width, uint32_t *out) {
    int n = 0;
    if (nodes[b].count > 0) {  // a leaf root: one child, the leaf itself
        out[n++] = b;
        return n;
    }
    out[n++] = nodes[b].first;
    out[n++] = nodes[b].first + 1;
    while (n < width) {
        int best = -1;
        float best_area = -1.0f;
        for (int i = 0; i < n; i++) {
            const lr_bvh_node &c = nodes[out[i]];
            if (c.count > 0) continue;
            float a = node_area(c);
            if (a > best_area) {
                best_area = a;
                best = i;
            }
        }
        if (best < 0) break;
        uint32_t opened = out[best];
        for (int i = n; i > best + 1; i--) out[i] = out[i - 1];
        out[best] = nodes[opened].first;
        out[best + 1] = nodes[opened].first + 1;
        n++;
    }
    return n;
}

template <int W> WideLayout collapse(const lr_bvh_node *nodes, size_t node_count) {
    WideLayout layout;
    layout.width = W;
    layout.node_bytes = padded_bytes<W>();
    std::vector<WideNode<W>> wide;
    std::vector<uint32_t> depth_pushes;  // stack bound below each wide node
    bool empty = node_count == 0 || nodes[0].lo[0] > nodes[0].hi[0];
    if (empty) {
        WideNode<W> w;
        std::memset(&w, 0, sizeof(w));
        for (int i = 0; i < W; i++) w.child[i] = 0xffffffffu;
        wide.push_back(w);
    } else {
        // Pre-order: each wide node's interior children follow it.
        struct Item {
            uint32_t binary;
            uint32_t wide_index;
        };
        std::vector<Item> stack;
        wide.emplace_back();
        stack.push_back(Item{0, 0});
        while (!stack.empty()) {
            Item it = stack.back();
            stack.pop_back();
            uint32_t kids[8];
            int n = gather(nodes, it.binary, W, kids);
            WideNode<W> w;
            std::memset(&w, 0, sizeof(w));
            std::vector<Item> children;
            for (int i = 0; i < W; i++) {
                if (i >= n) {
                    w.lo_x[i] = w.lo_y[i] = w.lo_z[i] = 1.0f;
                    w.hi_x[i] = w.hi_y[i] = w.hi_z[i] = -1.0f;
                    w.child[i] = 0xffffffffu;
                    w.count[i] = 0;
                    continue;
                }
                const lr_bvh_node &c = nodes[kids[i]];
                w.lo_x[i] = c.lo[0];
                w.lo_y[i] = c.lo[1];
                w.lo_z[i] = c.lo[2];
                w.hi_x[i] = c.hi[0];
                w.hi_y[i] = c.hi[1];
                w.hi_z[i] = c.hi[2];
                if (c.count > 0) {
                    w.child[i] = c.first;
                    w.count[i] = (uint8_t)c.count;
                } else {
                    uint32_t index = (uint32_t)wide.size();
                    wide.emplace_back();
                    w.child[i] = index;
                    w.count[i] = 0;
                    children.push_back(Item{kids[i], index});
                }
            }
            wide[it.wide_index] = w;
            for (auto c = children.rbegin(); c != children.rend(); ++c) stack.push_back(*c);
        }
    }
    // Stack bound: along any path, a node with k children pushes at most k - 1.
    std::vector<uint32_t> below(wide.size(), 0);
    for (size_t i = wide.size(); i-- > 0;) {
        const WideNode<W> &w = wide[i];
        int k = 0;
        uint32_t deepest = 0;
        for (int c = 0; c < W; c++) {
            if (w.child[c] == 0xffffffffu && w.count[c] == 0) continue;
            k++;
            if (w.count[c] == 0) deepest = std::max(deepest, below[w.child[c]]);
        }
        below[i] = (uint32_t)std::max(0, k - 1) + deepest;
    }
    layout.stack_bound = below.empty() ? 0 : below[0];
    if (W == 2) {
        // children follow their parents, so one backward sweep fills the ranges
        layout.ranges.assign(wide.size() * 2, 0);
        for (size_t i = wide.size(); i-- > 0;) {
            const WideNode<W> &w = wide[i];
            uint32_t lo = 0xffffffffu, hi = 0;
            for (int c = 0; c < W; c++) {
                if (w.child[c] == 0xffffffffu && w.count[c] == 0) continue;
                uint32_t a = w.count[c] ? w.child[c] : layout.ranges[w.child[c] * 2];
                uint32_t b = w.count[c] ? w.child[c] + w.count[c] - 1u : layout.ranges[w.child[c] * 2 + 1];
                lo = std::min(lo, a);
                hi = std::max(hi, b);
            }
            layout.ranges[i * 2] = lo;
            layout.ranges[i * 2 + 1] = hi;
        }
    }
    layout.nodes = wide.size();
    layout.bytes.assign(layout.nodes * layout.node_bytes, 0);
    for (size_t i = 0; i < wide.size(); i++) std::memcpy(&layout.bytes[i * layout.node_bytes], &wide[i], sizeof(WideNode<W>));
    return layout;
}

}  // namespace

WideLayout collapse_wide(const lr_bvh_node *nodes, size_t node_count, int width) {
    if (width == 2) return collapse<2>(nodes, node_count);
    if (width == 4) return collapse<4>(nodes, node_count);
    return collapse<8>(nodes, node_count);
}

}  // namespace lr
