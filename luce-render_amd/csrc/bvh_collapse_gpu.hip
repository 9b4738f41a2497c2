// Device versions of the upload-time repacking (bvh_wide.cpp, and the
// triangle packing of lr_scene_upload): width-W nodes (W = 2, 4, 8) collapsed
// from the reference's binary BVH, byte for byte the host collapse's layout,
// their stack bounds and the width-2 nodes' triangle slot ranges; and the
// leaf-order triangles as p0, e1, e2.
//
// The wide tree is discovered one wide level at a time: a thread gathers its
// wide node's children (the binary node's two children, then the
// largest-area interior one opened in place until W, ties to the first), and
// its interior children become the next level's nodes (temporary ids handed
// out one counter update per wave). The host collapse numbers wide nodes as
// it pops them from a stack: a node's interior children take the next
// indices when the node is visited, in pre-order. With k(v) the interior
// children of v and T(v) the sum of k over v's subtree (bottom-up), the
// first child of u sits at alloc(u), and its j-th interior child c_j at
// alloc(u) + j with alloc(c_j) = alloc(u) + k(u) + sum over i < j of
// T(c_i) (top-down).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "lr_internal.h"
#include "lr_scene.h"

namespace lr {
namespace gwide {

constexpr uint32_t kNone = 0xffffffffu;

__device__ __forceinline__ float node_area(const lr_bvh_node &n) {
    float x = n.hi[0] - n.lo[0], y = n.hi[1] - n.lo[1], z = n.hi[2] - n.lo[2];
    if (x < 0.0f || y < 0.0f || z < 0.0f) return 0.0f;
    return 2.0f * (x * y + y * z + z * x);
}

__device__ __forceinline__ uint64_t lanes_below(uint32_t lane) { return lane ? (~0ull >> (64u - lane)) : 0ull; }

// The binary descendants a wide node over binary node b holds (gather in
// bvh_wide.cpp).
template <int W> __device__ __forceinline__ int gather(const lr_bvh_node *nodes, uint32_t b, uint32_t out[W]) {
    const lr_bvh_node root = nodes[b];
    if (root.count > 0) {
        out[0] = b;
        return 1;
    }
    int n = 2;
    out[0] = root.first;
    out[1] = root.first + 1;
#pragma unroll
    for (int step = 2; step < W; step++) {
        int best = -1;
        float best_area = -1.0f;
#pragma unroll
        for (int i = 0; i < W; i++) {
            if (i >= n) break;
            const lr_bvh_node c = nodes[out[i]];
            if (c.count > 0) continue;
            float a = node_area(c);
            if (a > best_area) {
                best_area = a;
                best = i;
            }
        }
        if (best < 0) break;
        uint32_t opened = out[best];
        uint32_t f = nodes[opened].first;
#pragma unroll
        for (int i = W - 1; i > 0; i--)
            if (i > best + 1 && i <= n) out[i] = out[i - 1];
#pragma unroll
        for (int i = 0; i < W; i++) {
            if (i == best) out[i] = f;
            if (i == best + 1) out[i] = f + 1;
        }
        n++;
    }
    return n;
}

template <int W> struct Temp {
    uint32_t *binary;  // the binary node under each wide node
    uint32_t *kids;    // W binary children (kNone past the end)
    uint32_t *ktmp;    // W temporary ids of interior children
    uint32_t *kint;    // interior children
    uint32_t *kall;    // children
    uint32_t *tsum;    // T: interior children over the subtree
    uint32_t *below;   // stack bound below
    uint32_t *alloc;   // index of the first child
    uint32_t *index;   // final index
    uint2 *range;      // width 2: leaf-order slot range
};

template <int W>
__global__ __launch_bounds__(256) void expand_kernel(const lr_bvh_node *nodes, Temp<W> t, uint32_t begin, uint32_t end,
                                                     uint32_t *counter, uint32_t cap) {
    const uint32_t id = begin + blockIdx.x * 256u + threadIdx.x, lane = threadIdx.x & 63u;
    const bool live = id < end;
    uint32_t out[W];
    int n = 0;
    uint32_t k = 0;
    if (live) {
        n = gather<W>(nodes, t.binary[id], out);
        for (int i = 0; i < W; i++)
            if (i < n && nodes[out[i]].count == 0) k++;
    }
    // One counter update per wave for its interior children.
    uint32_t inc = k;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t v = __shfl_up(inc, o, 64);
        if ((int)lane >= o) inc += v;
    }
    uint32_t total = __shfl(inc, 63, 64), base = 0;
    if (lane == 0 && total) base = atomicAdd(counter, total);
    base = __shfl(base, 0, 64) + inc - k;
    if (!live) return;
    uint32_t j = 0;
    for (int i = 0; i < W; i++) {
        uint32_t kid = i < n ? out[i] : kNone, tmp = kNone;
        if (i < n && nodes[kid].count == 0) {
            tmp = base + j++;
            // a tree holds at most `cap` wide nodes; more (a malformed tree)
            // is reported by the host after the level
            if (tmp < cap) t.binary[tmp] = kid;
        }
        t.kids[(size_t)id * W + i] = kid;
        t.ktmp[(size_t)id * W + i] = tmp;
    }
    t.kint[id] = k;
    t.kall[id] = (uint32_t)n;
}

template <int W>
__global__ __launch_bounds__(256) void sums_kernel(const lr_bvh_node *nodes, Temp<W> t, uint32_t begin, uint32_t end) {
    const uint32_t id = begin + blockIdx.x * 256u + threadIdx.x;
    if (id >= end) return;
    uint32_t ts = t.kint[id], deepest = 0, lo = kNone, hi = 0;
    for (int i = 0; i < W; i++) {
        uint32_t c = t.ktmp[(size_t)id * W + i], kid = t.kids[(size_t)id * W + i];
        if (kid == kNone) continue;
        if (c != kNone) {
            ts += t.tsum[c];
            deepest = max(deepest, t.below[c]);
            if (W == 2) {
                lo = min(lo, t.range[c].x);
                hi = max(hi, t.range[c].y);
            }
        } else if (W == 2) {
            const lr_bvh_node leaf = nodes[kid];
            lo = min(lo, leaf.first);
            hi = max(hi, leaf.first + leaf.count - 1u);
        }
    }
    t.tsum[id] = ts;
    uint32_t kall = t.kall[id];
    t.below[id] = (kall > 0 ? kall - 1u : 0u) + deepest;
    if (W == 2) t.range[id] = make_uint2(lo, hi);
}

template <int W>
__global__ __launch_bounds__(256) void write_kernel(const lr_bvh_node *nodes, Temp<W> t, uint32_t begin, uint32_t end,
                                                    uint8_t *out, uint32_t node_bytes, uint2 *ranges,
                                                    uint32_t *slots) {
    const uint32_t id = begin + blockIdx.x * 256u + threadIdx.x;
    if (id >= end) return;
    const uint32_t a = t.alloc[id], at = t.index[id];
    uint32_t run = a + t.kint[id], j = 0;
    // Built in zeroed words so padding bytes are zero, as in the host layout.
    constexpr uint32_t kWords = (sizeof(WideNode<W>) + 3) / 4;
    union {
        WideNode<W> w;
        uint32_t u[kWords];
    } v;
    for (uint32_t q = 0; q < kWords; q++) v.u[q] = 0u;
    WideNode<W> &w = v.w;
    for (int i = 0; i < W; i++) {
        uint32_t kid = t.kids[(size_t)id * W + i], c = t.ktmp[(size_t)id * W + i];
        if (kid == kNone) {
            w.lo_x[i] = w.lo_y[i] = w.lo_z[i] = 1.0f;
            w.hi_x[i] = w.hi_y[i] = w.hi_z[i] = -1.0f;
            w.child[i] = kNone;
            w.count[i] = 0;
            continue;
        }
        const lr_bvh_node n = nodes[kid];
        w.lo_x[i] = n.lo[0];
        w.lo_y[i] = n.lo[1];
        w.lo_z[i] = n.lo[2];
        w.hi_x[i] = n.hi[0];
        w.hi_y[i] = n.hi[1];
        w.hi_z[i] = n.hi[2];
        if (c == kNone) {
            w.child[i] = n.first;
            w.count[i] = (uint8_t)n.count;
        } else {
            uint32_t ci = a + j++;
            t.index[c] = ci;
            t.alloc[c] = run;
            run += t.tsum[c];
            w.child[i] = ci;
            w.count[i] = 0;
        }
    }
    uint32_t *dst = reinterpret_cast<uint32_t *>(out + (size_t)at * node_bytes);
    for (uint32_t q = 0; q < kWords; q++) dst[q] = v.u[q];
    if (W == 2 && ranges) ranges[at] = t.range[id];
    if (slots)
        for (int i = 0; i < W; i++) slots[(size_t)at * W + i] = t.kids[(size_t)id * W + i];
}

// Leaf-order triangles: p0, e1 = p1 - p0, e2 = p2 - p0 (bvh.glsl:28-30), nine
// floats a slot.
__global__ __launch_bounds__(256) void pack_kernel(const float *positions, const uint32_t *indices, uint32_t n,
                                                  float *tris) {
    uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n) return;
    const float *p0 = positions + (size_t)indices[(size_t)t * 3] * 3;
    const float *p1 = positions + (size_t)indices[(size_t)t * 3 + 1] * 3;
    const float *p2 = positions + (size_t)indices[(size_t)t * 3 + 2] * 3;
    float *o = tris + (size_t)t * 9;
    o[0] = p0[0], o[1] = p0[1], o[2] = p0[2];
    o[3] = p1[0] - p0[0], o[4] = p1[1] - p0[1], o[5] = p1[2] - p0[2];
    o[6] = p2[0] - p0[0], o[7] = p2[1] - p0[1], o[8] = p2[2] - p0[2];
}

// Per-triangle rows of `words` 32-bit words gathered by the leaf order.
__global__ __launch_bounds__(256) void gather_rows_kernel(const uint32_t *src, const uint32_t *order, uint32_t n,
                                                         uint32_t words, uint32_t *dst) {
    uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= (uint64_t)n * words) return;
    uint64_t t = i / words, w = i % words;
    dst[i] = src[(uint64_t)order[t] * words + w];
}

template <class T> struct DeviceArray {
    T *p = nullptr;
    ~DeviceArray() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t count, const char *what) {
        hipError_t e = hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T));
        return e == hipSuccess ? 0 : hip_fail(e, what);
    }
};

template <int W>
int collapse(hipStream_t s, const lr_bvh_node *nodes, size_t node_count, DeviceLayout &layout, bool with_ranges,
             bool with_slots) {
    layout.node_bytes = (sizeof(WideNode<W>) + 63) / 64 * 64;
    hipError_t e;
    // Interior binary nodes, plus the root: a bound on the wide nodes.
    const size_t cap = node_count / 2 + 1;
    DeviceArray<uint32_t> binary, kids, ktmp, kint, kall, tsum, below, alloc, index, counter;
    DeviceArray<uint2> range;
    int rc;
    if ((rc = binary.alloc(cap, "hipMalloc(collapse)")) || (rc = kids.alloc(cap * W, "hipMalloc(collapse)")) ||
        (rc = ktmp.alloc(cap * W, "hipMalloc(collapse)")) || (rc = kint.alloc(cap, "hipMalloc(collapse)")) ||
        (rc = kall.alloc(cap, "hipMalloc(collapse)")) || (rc = tsum.alloc(cap, "hipMalloc(collapse)")) ||
        (rc = below.alloc(cap, "hipMalloc(collapse)")) || (rc = alloc.alloc(cap, "hipMalloc(collapse)")) ||
        (rc = index.alloc(cap, "hipMalloc(collapse)")) || (rc = counter.alloc(1, "hipMalloc(collapse)")) ||
        (W == 2 && (rc = range.alloc(cap, "hipMalloc(collapse)"))))
        return rc;
    Temp<W> t{binary.p, kids.p, ktmp.p, kint.p, kall.p, tsum.p, below.p, alloc.p, index.p, range.p};
    const uint32_t init[2] = {0u, 1u};  // root: binary node 0; one wide node handed out
    if ((e = hipMemcpyAsync(binary.p, &init[0], 4, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(counter.p, &init[1], 4, hipMemcpyHostToDevice, s)) != hipSuccess)
        return hip_fail(e, "hipMemcpy(collapse)");
    std::vector<std::pair<uint32_t, uint32_t>> levels{{0u, 1u}};
    for (;;) {
        auto [begin, end] = levels.back();
        hipLaunchKernelGGL(expand_kernel<W>, dim3((end - begin + 255) / 256), dim3(256), 0, s, nodes, t, begin, end,
                           counter.p, (uint32_t)cap);
        uint32_t used = 0;
        if ((e = hipMemcpyAsync(&used, counter.p, 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            return hip_fail(e, "collapse level");
        if (used > cap) return fail(LR_DEVICE_ERROR, "wide collapse found more nodes than the tree holds");
        if (used == end) break;
        if (levels.size() >= 4096) return fail(LR_LIMIT, "BVH deeper than 4096 levels");
        levels.push_back({end, used});
    }
    const uint32_t total = levels.back().second;
    for (size_t k = levels.size(); k-- > 0;) {
        auto [begin, end] = levels[k];
        hipLaunchKernelGGL(sums_kernel<W>, dim3((end - begin + 255) / 256), dim3(256), 0, s, nodes, t, begin, end);
    }
    layout.nodes = total;
    layout.bytes = (size_t)total * layout.node_bytes;
    if ((e = hipMalloc(&layout.data, layout.bytes)) != hipSuccess) return hip_fail(e, "hipMalloc(wide nodes)");
    if ((e = hipMemsetAsync(layout.data, 0, layout.bytes, s)) != hipSuccess) return hip_fail(e, "hipMemset(wide nodes)");
    if (with_ranges && W == 2) {
        if ((e = hipMalloc(&layout.ranges, (size_t)total * sizeof(uint2))) != hipSuccess)
            return hip_fail(e, "hipMalloc(node ranges)");
    }
    if (with_slots && (e = hipMalloc(&layout.slots, (size_t)total * W * 4)) != hipSuccess)
        return hip_fail(e, "hipMalloc(slot map)");
    const uint32_t root[2] = {0u, 1u};  // index 0, first child at 1
    if ((e = hipMemcpyAsync(index.p, &root[0], 4, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(alloc.p, &root[1], 4, hipMemcpyHostToDevice, s)) != hipSuccess)
        return hip_fail(e, "hipMemcpy(collapse)");
    for (auto [begin, end] : levels)
        hipLaunchKernelGGL(write_kernel<W>, dim3((end - begin + 255) / 256), dim3(256), 0, s, nodes, t, begin, end,
                           static_cast<uint8_t *>(layout.data), (uint32_t)layout.node_bytes,
                           static_cast<uint2 *>(layout.ranges), static_cast<uint32_t *>(layout.slots));
    uint32_t bound = 0;
    if ((e = hipMemcpyAsync(&bound, below.p, 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
        return hip_fail(e, "collapse");
    layout.stack_bound = bound;
    return 0;
}

}  // namespace gwide

int collapse_wide_device(lr_device *dev, const lr_bvh_node *nodes, size_t node_count, bool empty, int width,
                         DeviceLayout &layout, bool with_slots) {
    hipStream_t s = dev->stream;
    layout = DeviceLayout{};
    if (empty) {
        // One node with every child empty, as the host collapse writes it.
        WideLayout host = collapse_wide(nullptr, 0, width);
        hipError_t e;
        layout.node_bytes = host.node_bytes;
        layout.nodes = host.nodes;
        layout.bytes = host.bytes.size();
        layout.stack_bound = host.stack_bound;
        if ((e = hipMalloc(&layout.data, layout.bytes)) != hipSuccess ||
            (e = hipMemcpy(layout.data, host.bytes.data(), layout.bytes, hipMemcpyHostToDevice)) != hipSuccess)
            return hip_fail(e, "upload wide nodes");
        if (width == 2) {
            std::vector<uint32_t> r(2, 0u);
            if ((e = hipMalloc(&layout.ranges, 8)) != hipSuccess ||
                (e = hipMemcpy(layout.ranges, r.data(), 8, hipMemcpyHostToDevice)) != hipSuccess)
                return hip_fail(e, "upload node ranges");
        }
        return 0;
    }
    int rc = width == 2 ? gwide::collapse<2>(s, nodes, node_count, layout, true, with_slots)
             : width == 4 ? gwide::collapse<4>(s, nodes, node_count, layout, false, with_slots)
                          : gwide::collapse<8>(s, nodes, node_count, layout, false, with_slots);
    if (rc) {
        for (void *p : {layout.data, layout.ranges, layout.slots})
            if (p) (void)hipFree(p);
        layout = DeviceLayout{};
    }
    return rc;
}

void pack_triangles_device(lr_device *dev, const float *positions, const uint32_t *indices, size_t n, float *tris) {
    if (n == 0) return;
    hipLaunchKernelGGL(gwide::pack_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, dev->stream, positions,
                       indices, (uint32_t)n, tris);
}

}  // namespace lr

namespace lr {
void gather_rows_device(lr_device *dev, const uint32_t *src, const uint32_t *order, size_t n, uint32_t words,
                        uint32_t *dst) {
    if (n == 0) return;
    uint64_t items = (uint64_t)n * words;
    hipLaunchKernelGGL(gwide::gather_rows_kernel, dim3((uint32_t)((items + 255) / 256)), dim3(256), 0, dev->stream, src,
                       order, (uint32_t)n, words, dst);
}
}  // namespace lr

namespace lr {
namespace gwide {

// The quantized layout of a width-WD float layout (trace_device.h
// QuantNodeView), one thread a node. Per axis the grid starts at the node's
// low corner (the union of its children's boxes) with the smallest power-of-
// two step whose 255th line reaches the high corner; each child plane is
// rounded outward and then moved outward until origin + q * step, computed in
// f32 as the traversal decodes it, lies outside the float plane.
template <int WD> __global__ __launch_bounds__(256) void quantize_kernel(const uint32_t *src, uint32_t count, uint32_t *dst) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    constexpr int kSrc = WD == 4 ? 32 : 64;  // words a float node
    constexpr int kDst = WD == 4 ? 16 : 32;  // words a quantized node
    const uint32_t *n = src + (size_t)i * kSrc;
    uint32_t out[kDst];
#pragma unroll
    for (int k = 0; k < kDst; k++) out[k] = 0u;
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    bool any = false;
#pragma unroll
    for (int c = 0; c < WD; c++) {
        if (n[6 * WD + c] == kNone) continue;
        any = true;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo[a] = fminf(lo[a], __uint_as_float(n[2 * a * WD + c]));
            hi[a] = fmaxf(hi[a], __uint_as_float(n[(2 * a + 1) * WD + c]));
        }
    }
    float origin[3] = {0.0f, 0.0f, 0.0f}, step[3] = {1.0f, 1.0f, 1.0f};
    uint32_t exps = 0u;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int e = 127;
        if (any) {
            origin[a] = lo[a];
            float extent = hi[a] - origin[a];
            e = 1;
            if (extent > 0.0f) {
                int e2 = 0;
                frexpf(extent / 255.0f, &e2);
                e = min(max(e2 + 126, 1), 254);
            }
            while (e < 254 && origin[a] + 255.0f * __uint_as_float((uint32_t)e << 23) < hi[a]) e++;
        }
        step[a] = __uint_as_float((uint32_t)e << 23);
        exps |= (uint32_t)e << (8 * a);
        out[a] = __float_as_uint(origin[a]);
    }
    out[3] = exps;
    auto put_byte = [&](int at, uint32_t v) { out[at / 4] |= v << (8 * (at % 4)); };
#pragma unroll
    for (int c = 0; c < WD; c++) {
        uint32_t child = n[6 * WD + c];
        out[4 + 6 * WD / 4 + c] = child;
        put_byte(16 + 6 * WD + 4 * WD + c, (n[7 * WD + c / 4] >> (8 * (c % 4))) & 0xffu);
        if (child == kNone) continue;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float l = __uint_as_float(n[2 * a * WD + c]), h = __uint_as_float(n[(2 * a + 1) * WD + c]);
            float ql = fminf(fmaxf(floorf((l - origin[a]) / step[a]), 0.0f), 255.0f);
            float qh = fminf(fmaxf(ceilf((h - origin[a]) / step[a]), 0.0f), 255.0f);
            uint32_t q = (uint32_t)ql, r = (uint32_t)qh;
            while (q > 0u && origin[a] + (float)q * step[a] > l) q--;
            while (r < 255u && origin[a] + (float)r * step[a] < h) r++;
            put_byte(16 + 2 * a * WD + c, q);
            put_byte(16 + (2 * a + 1) * WD + c, r);
        }
    }
    uint32_t *d = dst + (size_t)i * kDst;
#pragma unroll
    for (int k = 0; k < kDst; k += 4) *reinterpret_cast<uint4 *>(d + k) = make_uint4(out[k], out[k + 1], out[k + 2], out[k + 3]);
}

}  // namespace gwide

int quantize_wide_device(lr_device *dev, const void *src, size_t count, int width, void **dst) {
    size_t node_bytes = width == 4 ? 64 : 128;
    size_t bytes = std::max<size_t>(count, 1) * node_bytes;
    hipError_t e;
    if (!*dst && (e = hipMalloc(dst, bytes)) != hipSuccess) return hip_fail(e, "hipMalloc(quantized nodes)");
    if (count == 0) {
        if ((e = hipMemsetAsync(*dst, 0, bytes, dev->stream)) != hipSuccess) return hip_fail(e, "hipMemset(quantized nodes)");
        return 0;
    }
    dim3 grid((uint32_t)((count + 255) / 256));
    if (width == 4)
        hipLaunchKernelGGL(gwide::quantize_kernel<4>, grid, dim3(256), 0, dev->stream, static_cast<const uint32_t *>(src),
                           (uint32_t)count, static_cast<uint32_t *>(*dst));
    else
        hipLaunchKernelGGL(gwide::quantize_kernel<8>, grid, dim3(256), 0, dev->stream, static_cast<const uint32_t *>(src),
                           (uint32_t)count, static_cast<uint32_t *>(*dst));
    e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, "quantize wide nodes");
}

}  // namespace lr
