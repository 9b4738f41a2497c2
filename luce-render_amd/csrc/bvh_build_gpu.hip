// Device build of the reference's binned-SAH binary BVH (src/render/bvh.lucb:
// 66-217): the same tree, node numbering and leaf order as the host builder
// (bvh_build.cpp), built level by level on the GPU.
//
// A level's nodes are processed one wave each (nodes larger than kBigNode by
// several workgroups each, the chunked kernels below): the node's box and
// centroid bounds (min/max reductions, exact in any order), 12 centroid bins
// per axis (counts and boxes by LDS atomics on order-preserving integer
// encodings of the floats), the SAH costs in the reference's f32 operation
// order, and the split decision of bvh.lucb:112-120. The partition is the
// reference's in-place swap loop (bvh.lucb:185-196) computed in closed form:
// with L the node's left count, a left element in [0, L) stays; the k-th right
// element in [0, L) goes to j(k-1) - 1, where j(r) is the position of the r-th
// left element of [L, n) counted from the end and j(0) = n; that left element
// goes to the r-th right's place; right elements past L shift down by one and
// the one at L (if right) goes to j(m) - 1, m the number of rights in [0, L).
// Two table sweeps give every element its destination from its rank among
// the lefts before it.
//
// Child pairs take temporary ids from a counter, so a level's nodes are one
// contiguous id range. Afterwards two passes over the levels renumber them in
// the reference's order (a split node's pair takes the next two indices when
// the node is visited, pre-order, left first): interior counts bottom-up, then
// each node's pre-order rank among interior nodes top-down; the pair of the
// interior node of rank r is 1 + 2r.
//
// Node boxes equal the host builder's as values; where a box plane is a zero
// that occurs with both signs, the sign the reductions keep may differ (the
// sequential builder keeps the later one). Bins, costs and traversal results
// do not depend on that sign.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "lr_internal.h"
#include "lr_scene.h"

namespace lr {
namespace gbvh {

constexpr int kBins = 12;          // bvh.lucb:29
constexpr uint32_t kLeafSize = 4;  // bvh.lucb:30
constexpr uint32_t kLeafLimit = 16;  // bvh.lucb:32
constexpr uint32_t kChunk = 4096;    // positions per workgroup in the chunked kernels
constexpr uint32_t kBigNode = 16384; // nodes above this go to the chunked kernels
constexpr uint32_t kSlots = 3 * kBins;  // (axis, bin)

__device__ __forceinline__ float mn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float mx(float a, float b) { return a > b ? a : b; }
// Order-preserving map of floats to unsigned integers (and back).
__device__ __forceinline__ uint32_t enc(float f) {
    uint32_t u = __float_as_uint(f);
    return u ^ ((u >> 31) ? 0xffffffffu : 0x80000000u);
}
__device__ __forceinline__ float dec(uint32_t u) { return __uint_as_float(u ^ ((u >> 31) ? 0x80000000u : 0xffffffffu)); }

__device__ __forceinline__ float box_area(const float lo[3], const float hi[3]) {
    float x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
    if (x < 0.0f || y < 0.0f || z < 0.0f) return 0.0f;
    float s = x * y + y * z;
    s = s + z * x;
    return 2.0f * s;
}

// bin_of (bvh.lucb:136-145) on one axis of the centroid bounds.
__device__ __forceinline__ int bin_of(float c, float lo, float hi) {
    float extent = hi - lo;
    if (extent <= 0.0f) return 0;
    float q = ((float)kBins * (c - lo)) / extent;
    int64_t at = (int64_t)q;
    if (at < 0) return 0;
    return at < kBins ? (int)at : kBins - 1;
}

__device__ __forceinline__ float centroid(float lo, float hi) { return 0.5f * (lo + hi); }

__device__ __forceinline__ uint64_t lanes_below(uint32_t lane) { return lane ? (~0ull >> (64u - lane)) : 0ull; }

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = mn(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = mx(v, __shfl_xor(v, o, 64));
    return v;
}

// Triangle boxes (the reference grows an empty box by the three corners) and
// the identity order.
__global__ __launch_bounds__(256) void prep_kernel(const float *positions, const uint32_t *indices, uint32_t n,
                                                  float4 *blo, float4 *bhi, uint32_t *order) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (int k = 0; k < 3; k++) {
        const float *p = positions + (size_t)indices[(size_t)t * 3 + k] * 3;
        for (int d = 0; d < 3; d++) {
            lo[d] = mn(lo[d], p[d]);
            hi[d] = mx(hi[d], p[d]);
        }
    }
    blo[t] = make_float4(lo[0], lo[1], lo[2], 0.0f);
    bhi[t] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    order[t] = t;
}

struct Bounds {
    float lo[3], hi[3];   // the node's box
    float ml[3], mh[3];   // its centroid bounds
};

struct LevelArgs {
    const float4 *blo, *bhi;
    const uint32_t *cur;  // this level's order (segments of its nodes)
    uint32_t *next;       // the next level's order
    uint32_t *fin;        // the final leaf order
    uint32_t *tab;        // partition tables, one word per position
    lr_bvh_node *nodes;   // temporary nodes: pending {first, count}, then final
    uint32_t *counter;    // temporary ids handed out
    uint32_t begin, end;  // this level's ids
    uint32_t big;         // nodes above this many triangles are done by the chunked kernels
};

// The best split over gathered bins (best_split, bvh.lucb:147-182) by one wave:
// lane axis * 11 + (s - 1) prices the split before bin s on `axis`; the
// lowest-priced lane wins, the lowest lane on a tie (the reference's strict <
// in axis-then-bin order).
__device__ __forceinline__ void best_split(const Bounds &b, const uint32_t *cnt, const uint32_t *bl, const uint32_t *bh,
                                           uint32_t lane, int &axis_out, int &bin_out, float &cost_out) {
    float parent = box_area(b.lo, b.hi);
    bool valid = false;
    float cost = 3.0e38f;
    if (lane < 33u && parent > 0.0f) {
        int axis = (int)(lane / 11u), s = (int)(lane % 11u) + 1;
        if (b.mh[axis] - b.ml[axis] > 0.0f) {
            float llo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, lhi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
            float rlo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, rhi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
            uint32_t lc = 0, rc = 0;
            for (int k = 0; k < kBins; k++) {
                uint32_t slot = (uint32_t)(axis * kBins + k);
                if (k < s) {
                    lc += cnt[slot];
                    for (int d = 0; d < 3; d++) {
                        llo[d] = mn(llo[d], dec(bl[slot * 3 + d]));
                        lhi[d] = mx(lhi[d], dec(bh[slot * 3 + d]));
                    }
                } else {
                    rc += cnt[slot];
                    for (int d = 0; d < 3; d++) {
                        rlo[d] = mn(rlo[d], dec(bl[slot * 3 + d]));
                        rhi[d] = mx(rhi[d], dec(bh[slot * 3 + d]));
                    }
                }
            }
            if (lc && rc) {
                float l = box_area(llo, lhi) * (float)lc;
                float r = box_area(rlo, rhi) * (float)rc;
                cost = 1.0f + (l + r) / parent;
                valid = cost < 3.0e38f;
            }
        }
    }
    float key = valid ? cost : __builtin_inff();
    float best = key;
    for (int o = 32; o > 0; o >>= 1) best = fminf(best, __shfl_xor(best, o, 64));
    uint64_t winners = __ballot(valid && key == best);
    if (!winners) {
        axis_out = -1;
        bin_out = 0;
        cost_out = 3.0e38f;
        return;
    }
    uint32_t w = (uint32_t)__ffsll((long long)winners) - 1u;
    axis_out = (int)(w / 11u);
    bin_out = (int)(w % 11u) + 1;
    cost_out = best;
}

// The destination of position i (closed form of the swap partition) given
// whether it goes left and the lefts before it; `tab` the node's tables.
// Every table index lies in [0, n) when L counts the node's lefts (the bins
// and the sweeps classify with the same arithmetic); the checks keep a
// violation inside the node's segment.
__device__ __forceinline__ uint32_t table_at(const uint32_t *tab, uint32_t at, uint32_t n) {
    return tab[at < n ? at : n - 1u];
}
__device__ __forceinline__ uint32_t destination(bool left, uint32_t i, uint32_t lb, uint32_t L, uint32_t n,
                                                const uint32_t *tab) {
    if (left) return i < L ? i : table_at(tab, L - lb - 1u, n);
    if (i < L) {
        uint32_t k = i - lb + 1u;
        return (k == 1u ? n : table_at(tab, L + k - 2u, n)) - 1u;
    }
    if (i > L) return i - 1u;
    uint32_t m = L - lb;
    return (m == 0u ? n : table_at(tab, L + m - 1u, n)) - 1u;
}

// The tables of position i: the k-th right of [0, L) records its place at
// tab[k - 1]; the r-th left of [L, n) from the end at tab[L + r - 1].
__device__ __forceinline__ void table_entry(bool left, uint32_t i, uint32_t lb, uint32_t L, uint32_t n, uint32_t *tab) {
    uint32_t at = n;
    if (i < L && !left) at = i - lb;
    else if (i >= L && left) at = L + (L - lb) - 1u;
    if (at < n) tab[at] = i;
}

__global__ __launch_bounds__(64) void level_kernel(LevelArgs a) {
    __shared__ uint32_t cnt[kSlots];
    __shared__ uint32_t bl[kSlots * 3], bh[kSlots * 3];
    const uint32_t lane = threadIdx.x;
    for (uint32_t id = a.begin + blockIdx.x; id < a.end; id += gridDim.x) {
        const lr_bvh_node pending = a.nodes[id];
        const uint32_t first = pending.first, count = pending.count;
        if (count > a.big) continue;
        const uint32_t *cur = a.cur + first;
        // The node's box and centroid bounds.
        Bounds b;
        for (int d = 0; d < 3; d++) b.lo[d] = b.ml[d] = 3.0e38f, b.hi[d] = b.mh[d] = -3.0e38f;
        for (uint32_t i = lane; i < count; i += 64u) {
            uint32_t p = cur[i];
            float4 l = a.blo[p], h = a.bhi[p];
            float lv[3] = {l.x, l.y, l.z}, hv[3] = {h.x, h.y, h.z};
            for (int d = 0; d < 3; d++) {
                b.lo[d] = mn(b.lo[d], lv[d]);
                b.hi[d] = mx(b.hi[d], hv[d]);
                float c = centroid(lv[d], hv[d]);
                b.ml[d] = mn(b.ml[d], c);
                b.mh[d] = mx(b.mh[d], c);
            }
        }
        for (int d = 0; d < 3; d++) {
            b.lo[d] = wave_min(b.lo[d]);
            b.ml[d] = wave_min(b.ml[d]);
            b.hi[d] = wave_max(b.hi[d]);
            b.mh[d] = wave_max(b.mh[d]);
        }
        int axis = -1, bin = 0;
        float cost = 3.0e38f;
        if (count > kLeafSize) {
            for (uint32_t j = lane; j < kSlots; j += 64u) cnt[j] = 0u;
            for (uint32_t j = lane; j < kSlots * 3u; j += 64u) {
                bl[j] = enc(3.0e38f);
                bh[j] = enc(-3.0e38f);
            }
            __syncthreads();
            for (uint32_t i = lane; i < count; i += 64u) {
                uint32_t p = cur[i];
                float4 l = a.blo[p], h = a.bhi[p];
                float lv[3] = {l.x, l.y, l.z}, hv[3] = {h.x, h.y, h.z};
                for (int ax = 0; ax < 3; ax++) {
                    uint32_t slot = (uint32_t)(ax * kBins + bin_of(centroid(lv[ax], hv[ax]), b.ml[ax], b.mh[ax]));
                    atomicAdd(&cnt[slot], 1u);
                    for (int d = 0; d < 3; d++) {
                        atomicMin(&bl[slot * 3 + d], enc(lv[d]));
                        atomicMax(&bh[slot * 3 + d], enc(hv[d]));
                    }
                }
            }
            __syncthreads();
            best_split(b, cnt, bl, bh, lane, axis, bin, cost);
        }
        uint32_t left_count = 0;
        bool split = count > kLeafSize && (cost < (float)count || count > kLeafLimit);
        if (split && axis >= 0) {
            for (int k = 0; k < bin; k++) left_count += cnt[axis * kBins + k];
            const uint32_t L = left_count;
            uint32_t *tab = a.tab + first;
            uint32_t base = 0;
            for (uint32_t i0 = 0; i0 < count; i0 += 64u) {
                uint32_t i = i0 + lane;
                bool left = false;
                if (i < count) {
                    uint32_t p = cur[i];
                    float lo = (&a.blo[p].x)[axis], hi = (&a.bhi[p].x)[axis];
                    left = bin_of(centroid(lo, hi), b.ml[axis], b.mh[axis]) < bin;
                }
                uint64_t m = __ballot(left);
                uint32_t lb = base + (uint32_t)__popcll(m & lanes_below(lane));
                if (i < count) table_entry(left, i, lb, L, count, tab);
                base += (uint32_t)__popcll(m);
            }
            __threadfence_block();
            __syncthreads();
            base = 0;
            for (uint32_t i0 = 0; i0 < count; i0 += 64u) {
                uint32_t i = i0 + lane;
                bool left = false;
                uint32_t p = 0;
                if (i < count) {
                    p = cur[i];
                    float lo = (&a.blo[p].x)[axis], hi = (&a.bhi[p].x)[axis];
                    left = bin_of(centroid(lo, hi), b.ml[axis], b.mh[axis]) < bin;
                }
                uint64_t m = __ballot(left);
                uint32_t lb = base + (uint32_t)__popcll(m & lanes_below(lane));
                if (i < count) {
                    uint32_t d = destination(left, i, lb, L, count, tab);
                    if (d < count) a.next[first + d] = p;
                }
                base += (uint32_t)__popcll(m);
            }
            if (left_count == 0 || left_count == count) left_count = count / 2u;
        } else if (split) {
            // No separating bin: halve by count, the order as it is.
            for (uint32_t i = lane; i < count; i += 64u) a.next[first + i] = cur[i];
            left_count = count / 2u;
        }
        if (!split) {
            for (uint32_t i = lane; i < count; i += 64u) a.fin[first + i] = cur[i];
            if (lane == 0) a.nodes[id] = lr_bvh_node{{b.lo[0], b.lo[1], b.lo[2]}, first, {b.hi[0], b.hi[1], b.hi[2]}, count};
        } else if (lane == 0) {
            uint32_t pair = atomicAdd(a.counter, 2u);
            a.nodes[id] = lr_bvh_node{{b.lo[0], b.lo[1], b.lo[2]}, pair, {b.hi[0], b.hi[1], b.hi[2]}, 0u};
            a.nodes[pair] = lr_bvh_node{{0, 0, 0}, first, {0, 0, 0}, left_count};
            a.nodes[pair + 1] = lr_bvh_node{{0, 0, 0}, first + left_count, {0, 0, 0}, count - left_count};
        }
        __syncthreads();  // the LDS bins are reused by the next node
    }
}

// ------------------------------------------------- renumbering in the reference's order

// Interior nodes in each subtree (bottom-up, one level at a time).
__global__ __launch_bounds__(256) void interior_kernel(const lr_bvh_node *nodes, uint32_t *inner, uint32_t begin,
                                                      uint32_t end) {
    uint32_t id = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= end) return;
    const lr_bvh_node n = nodes[id];
    inner[id] = n.count ? 0u : 1u + inner[n.first] + inner[n.first + 1];
}

// Top-down: a node's final index and its pre-order rank among interior nodes
// are set by its parent; an interior node of rank r has its pair at 1 + 2r.
__global__ __launch_bounds__(256) void emit_kernel(const lr_bvh_node *nodes, const uint32_t *inner, uint32_t *rank,
                                                  uint32_t *index, lr_bvh_node *out, uint32_t begin, uint32_t end) {
    uint32_t id = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= end) return;
    lr_bvh_node n = nodes[id];
    uint32_t at = index[id];
    if (n.count == 0) {
        uint32_t r = rank[id], c = n.first, pair = 1u + 2u * r;
        index[c] = pair;
        index[c + 1] = pair + 1u;
        rank[c] = r + 1u;
        rank[c + 1] = r + 1u + inner[c];
        n.first = pair;
    }
    out[at] = n;
}

}  // namespace gbvh

namespace {

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

}  // namespace

// Builds the tree of `n` triangles whose points and indices are on the device;
// writes `node_count` nodes to `out` (room for max(1, 2n)) and the leaf order.
int build_bvh_device(lr_device *dev, const float *positions, const uint32_t *indices, size_t n, lr_bvh_node *out,
                     uint32_t *order, size_t *node_count) {
    using namespace gbvh;
    hipStream_t s = dev->stream;
    hipError_t e;
    if (n == 0) {
        lr_bvh_node empty{{1.0f, 1.0f, 1.0f}, 0, {-1.0f, -1.0f, -1.0f}, 0};
        if ((e = hipMemcpyAsync(out, &empty, sizeof(empty), hipMemcpyHostToDevice, s)) != hipSuccess)
            return hip_fail(e, "hipMemcpy(empty tree)");
        *node_count = 1;
        return 0;
    }
    DeviceArray<float4> blo, bhi;
    DeviceArray<uint32_t> ord[2], tab, counter, inner, rank, index;
    DeviceArray<lr_bvh_node> tmp;
    int rc;
    const size_t cap = 2 * n;
    if ((rc = blo.alloc(n, "hipMalloc(build boxes)")) || (rc = bhi.alloc(n, "hipMalloc(build boxes)")) ||
        (rc = ord[0].alloc(n, "hipMalloc(build order)")) || (rc = ord[1].alloc(n, "hipMalloc(build order)")) ||
        (rc = tab.alloc(n, "hipMalloc(build tables)")) || (rc = counter.alloc(1, "hipMalloc(build counter)")) ||
        (rc = tmp.alloc(cap, "hipMalloc(build nodes)")))
        return rc;
    const uint32_t n32 = (uint32_t)n;
    hipLaunchKernelGGL(prep_kernel, dim3((n32 + 255) / 256), dim3(256), 0, s, positions, indices, n32, blo.p, bhi.p,
                       ord[0].p);
    lr_bvh_node root{{0, 0, 0}, 0, {0, 0, 0}, n32};
    uint32_t one = 1;
    if ((e = hipMemcpyAsync(tmp.p, &root, sizeof(root), hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(counter.p, &one, 4, hipMemcpyHostToDevice, s)) != hipSuccess)
        return hip_fail(e, "hipMemcpy(build root)");
    std::vector<std::pair<uint32_t, uint32_t>> levels{{0u, 1u}};
    int c = 0;
    for (;;) {
        auto [begin, end] = levels.back();
        LevelArgs a{blo.p, bhi.p, ord[c].p, ord[c ^ 1].p, order, tab.p, tmp.p, counter.p, begin, end, 0xffffffffu};
        uint32_t blocks = std::min<uint32_t>(end - begin, 65536u);
        hipLaunchKernelGGL(level_kernel, dim3(blocks), dim3(64), 0, s, a);
        uint32_t used = 0;
        if ((e = hipMemcpyAsync(&used, counter.p, 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            return hip_fail(e, "BVH build level");
        if (used > cap) return fail(LR_DEVICE_ERROR, "BVH build handed out more nodes than the tree can hold");
        if (used == end) break;
        if (levels.size() >= 4096) return fail(LR_LIMIT, "BVH deeper than 4096 levels");
        levels.push_back({end, used});
        c ^= 1;
    }
    const uint32_t total = levels.back().second;
    if ((rc = inner.alloc(total, "hipMalloc(build renumber)")) || (rc = rank.alloc(total, "hipMalloc(build renumber)")) ||
        (rc = index.alloc(total, "hipMalloc(build renumber)")))
        return rc;
    for (size_t k = levels.size(); k-- > 0;) {
        auto [begin, end] = levels[k];
        hipLaunchKernelGGL(interior_kernel, dim3((end - begin + 255) / 256), dim3(256), 0, s, tmp.p, inner.p, begin, end);
    }
    uint32_t zero[1] = {0};
    if ((e = hipMemcpyAsync(rank.p, zero, 4, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(index.p, zero, 4, hipMemcpyHostToDevice, s)) != hipSuccess)
        return hip_fail(e, "hipMemcpy(build renumber)");
    for (auto [begin, end] : levels)
        hipLaunchKernelGGL(emit_kernel, dim3((end - begin + 255) / 256), dim3(256), 0, s, tmp.p, inner.p, rank.p,
                           index.p, out, begin, end);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "BVH build launch");
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "BVH build");
    *node_count = total;
    return 0;
}

}  // namespace lr

using namespace lr;

extern "C" int lr_bvh_build_device(lr_device *device, const float *positions, size_t points, const uint32_t *indices,
                                   size_t triangles, lr_bvh_node *nodes, size_t *node_count, uint32_t *order) {
    if (!device || !nodes || !node_count || (triangles && (!positions || !indices || !order)))
        return fail(LR_INVALID, "lr_bvh_build_device needs a device, positions, indices, nodes and order");
    if (points > 4294967295ull) return fail(LR_LIMIT, "a scene holds at most 2^32 points");
    if (triangles >= (1ull << 27)) return fail(LR_LIMIT, "a scene holds at most 2^27 triangles");
    for (size_t i = 0; i < triangles * 3; i++)
        if (indices[i] >= points) return fail(LR_INVALID, "a triangle names a point the scene lacks");
    (void)hipSetDevice(device->ordinal);
    hipStream_t s = device->stream;
    float *dpos = nullptr;
    uint32_t *didx = nullptr, *dorder = nullptr;
    lr_bvh_node *dnodes = nullptr;
    auto release = [&] {
        for (void *p : {(void *)dpos, (void *)didx, (void *)dorder, (void *)dnodes})
            if (p) (void)hipFree(p);
    };
    hipError_t e;
    size_t cap = std::max<size_t>(1, 2 * triangles);
    if ((e = hipMalloc((void **)&dpos, std::max<size_t>(points, 1) * 12)) != hipSuccess ||
        (e = hipMalloc((void **)&didx, std::max<size_t>(triangles, 1) * 12)) != hipSuccess ||
        (e = hipMalloc((void **)&dorder, std::max<size_t>(triangles, 1) * 4)) != hipSuccess ||
        (e = hipMalloc((void **)&dnodes, cap * sizeof(lr_bvh_node))) != hipSuccess) {
        release();
        return hip_fail(e, "hipMalloc(BVH build)");
    }
    if (triangles && ((e = hipMemcpyAsync(dpos, positions, points * 12, hipMemcpyHostToDevice, s)) != hipSuccess ||
                      (e = hipMemcpyAsync(didx, indices, triangles * 12, hipMemcpyHostToDevice, s)) != hipSuccess)) {
        release();
        return hip_fail(e, "hipMemcpy(BVH build inputs)");
    }
    size_t count = 0;
    int rc = build_bvh_device(device, dpos, didx, triangles, dnodes, dorder, &count);
    if (!rc && ((e = hipMemcpyAsync(nodes, dnodes, count * sizeof(lr_bvh_node), hipMemcpyDeviceToHost, s)) != hipSuccess ||
                (triangles && (e = hipMemcpyAsync(order, dorder, triangles * 4, hipMemcpyDeviceToHost, s)) != hipSuccess) ||
                (e = hipStreamSynchronize(s)) != hipSuccess))
        rc = hip_fail(e, "hipMemcpy(BVH build results)");
    release();
    if (!rc) *node_count = count;
    return rc;
}
