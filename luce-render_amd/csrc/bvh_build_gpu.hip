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
constexpr uint32_t kChunk = 2048;    // positions per workgroup in the chunked kernels
constexpr uint32_t kBigNode = 2048;  // nodes above this go to the chunked kernels
constexpr uint32_t kSlots = 3 * kBins;  // (axis, bin)
constexpr uint32_t kReplicas = 16;      // copies of a large node's bounds and bins the chunks merge into

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
    uint32_t batch;       // nodes per wave (at most 64)
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

// Bins the triangles at positions [i0, i1) of one thread's contiguous run:
// neighbours in the order tend to share a bin, so each axis keeps a running
// bin (count and box) and adds it to the LDS bins only when the bin changes,
// instead of every lane of a wave hitting the same LDS words at once.
__device__ __forceinline__ void bin_run(const uint32_t *cur, uint32_t i0, uint32_t i1, const float4 *blo,
                                        const float4 *bhi, const Bounds &b, uint32_t *cnt, uint32_t *bl,
                                        uint32_t *bh) {
    int rb[3] = {-1, -1, -1};
    uint32_t rc[3] = {0u, 0u, 0u};
    float rlo[3][3], rhi[3][3];
    auto flush = [&](int ax) {
        if (!rc[ax]) return;
        uint32_t slot = (uint32_t)(ax * kBins + rb[ax]);
        atomicAdd(&cnt[slot], rc[ax]);
        for (int d = 0; d < 3; d++) {
            atomicMin(&bl[slot * 3 + d], enc(rlo[ax][d]));
            atomicMax(&bh[slot * 3 + d], enc(rhi[ax][d]));
        }
    };
    for (uint32_t i = i0; i < i1; i++) {
        uint32_t p = cur[i];
        float4 l = blo[p], h = bhi[p];
        float lv[3] = {l.x, l.y, l.z}, hv[3] = {h.x, h.y, h.z};
        for (int ax = 0; ax < 3; ax++) {
            int bn = bin_of(centroid(lv[ax], hv[ax]), b.ml[ax], b.mh[ax]);
            if (bn != rb[ax]) {
                flush(ax);
                rb[ax] = bn;
                rc[ax] = 0u;
                for (int d = 0; d < 3; d++) rlo[ax][d] = 3.0e38f, rhi[ax][d] = -3.0e38f;
            }
            rc[ax]++;
            for (int d = 0; d < 3; d++) {
                rlo[ax][d] = mn(rlo[ax][d], lv[d]);
                rhi[ax][d] = mx(rhi[ax][d], hv[d]);
            }
        }
    }
    for (int ax = 0; ax < 3; ax++) flush(ax);
}

__global__ __launch_bounds__(64) void level_kernel(LevelArgs a) {
    __shared__ uint32_t cnt[kSlots];
    __shared__ uint32_t bl[kSlots * 3], bh[kSlots * 3];
    const uint32_t lane = threadIdx.x;
    // Batches of up to 64 nodes: lane j keeps node j's split, so the batch's
    // child pairs take one counter update (one counter for the whole level
    // would otherwise serialize an atomic per node).
    for (uint32_t batch = a.begin + blockIdx.x * a.batch; batch < a.end; batch += gridDim.x * a.batch) {
      const uint32_t nb = min(a.batch, a.end - batch);
      bool my_split = false;
      uint32_t my_first = 0, my_count = 0, my_left = 0;
      float my_lo[3] = {0.0f, 0.0f, 0.0f}, my_hi[3] = {0.0f, 0.0f, 0.0f};
      for (uint32_t j = 0; j < nb; j++) {
        const uint32_t id = batch + j;
        const lr_bvh_node pending = a.nodes[id];
        const uint32_t first = pending.first, count = pending.count;
        // count 0: a large node the chunked kernels have split already
        if (count == 0 || count > a.big) continue;
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
            const uint32_t run = (count + 63u) / 64u, r0 = min(count, lane * run);
            bin_run(cur, r0, min(count, r0 + run), a.blo, a.bhi, b, cnt, bl, bh);
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
        } else if (lane == j) {
            my_split = true;
            my_first = first;
            my_count = count;
            my_left = left_count;
            for (int d = 0; d < 3; d++) my_lo[d] = b.lo[d], my_hi[d] = b.hi[d];
        }
        __syncthreads();  // the LDS bins are reused by the next node
      }
      const uint64_t m = __ballot(my_split);
      if (!m) continue;
      uint32_t pair = 0;
      if (lane == 0) pair = atomicAdd(a.counter, 2u * (uint32_t)__popcll(m));
      pair = __shfl(pair, 0, 64) + 2u * (uint32_t)__popcll(m & lanes_below(lane));
      if (my_split) {
          a.nodes[batch + lane] = lr_bvh_node{{my_lo[0], my_lo[1], my_lo[2]}, pair, {my_hi[0], my_hi[1], my_hi[2]}, 0u};
          a.nodes[pair] = lr_bvh_node{{0, 0, 0}, my_first, {0, 0, 0}, my_left};
          a.nodes[pair + 1] = lr_bvh_node{{0, 0, 0}, my_first + my_left, {0, 0, 0}, my_count - my_left};
      }
    }
}

// ------------------------------------------------- large nodes: several workgroups each

// A node above kBigNode triangles, its run of chunks, and its gathered state:
// box and centroid bounds, bins (both as encoded floats) and the split.
struct BigNode {
    uint32_t id, first, count, chunk0, chunks;
};
struct Chunk {
    uint32_t big, begin, end;  // positions [begin, end) of big node `big`
};
struct BigState {
    uint32_t box[12];  // lo xyz, hi xyz, centroid lo xyz, centroid hi xyz
    uint32_t cnt[kSlots];
    uint32_t bl[kSlots * 3], bh[kSlots * 3];
    int axis, bin;
    uint32_t left, halve;
};

struct BigArgs {
    const float4 *blo, *bhi;
    const uint32_t *cur;
    uint32_t *next, *tab;
    lr_bvh_node *nodes;
    uint32_t *counter;
    const BigNode *bigs;
    const Chunk *chunks;
    BigState *st;
    uint32_t *chunk_left, *chunk_base;
    BigNode *next_bigs;
    uint32_t *next_big_count;
};

__device__ __forceinline__ Bounds big_bounds(const BigState &x) {
    Bounds b;
    for (int d = 0; d < 3; d++) {
        b.lo[d] = dec(x.box[d]);
        b.hi[d] = dec(x.box[3 + d]);
        b.ml[d] = dec(x.box[6 + d]);
        b.mh[d] = dec(x.box[9 + d]);
    }
    return b;
}

// The exclusive rank of `flag` among a 256-thread round's flags, and the
// round's total.
__device__ __forceinline__ uint32_t block_prefix(bool flag, uint32_t *waves, uint32_t &total) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint64_t m = __ballot(flag);
    if (lane == 0) waves[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (uint32_t w = 0; w < 4u; w++) {
        uint32_t c = waves[w];
        before += w < wave ? c : 0u;
        total += c;
    }
    __syncthreads();
    return before + (uint32_t)__popcll(m & lanes_below(lane));
}

// st holds kReplicas states per large node; replica 0 receives the merge.
__global__ __launch_bounds__(64) void big_init_kernel(BigState *st) {
    BigState &x = st[blockIdx.x];
    const uint32_t t = threadIdx.x;
    if (t < 12u) x.box[t] = enc((t < 3u || (t >= 6u && t < 9u)) ? 3.0e38f : -3.0e38f);
    for (uint32_t j = t; j < kSlots; j += 64u) x.cnt[j] = 0u;
    for (uint32_t j = t; j < kSlots * 3u; j += 64u) {
        x.bl[j] = enc(3.0e38f);
        x.bh[j] = enc(-3.0e38f);
    }
}

__global__ __launch_bounds__(256) void big_bounds_kernel(BigArgs a) {
    const Chunk ch = a.chunks[blockIdx.x];
    float v[12];
    for (int d = 0; d < 3; d++) v[d] = v[6 + d] = 3.0e38f, v[3 + d] = v[9 + d] = -3.0e38f;
    for (uint32_t i = ch.begin + threadIdx.x; i < ch.end; i += 256u) {
        uint32_t p = a.cur[i];
        float4 l = a.blo[p], h = a.bhi[p];
        float lv[3] = {l.x, l.y, l.z}, hv[3] = {h.x, h.y, h.z};
        for (int d = 0; d < 3; d++) {
            v[d] = mn(v[d], lv[d]);
            v[3 + d] = mx(v[3 + d], hv[d]);
            float c = centroid(lv[d], hv[d]);
            v[6 + d] = mn(v[6 + d], c);
            v[9 + d] = mx(v[9 + d], c);
        }
    }
    for (int d = 0; d < 3; d++) {
        v[d] = wave_min(v[d]);
        v[3 + d] = wave_max(v[3 + d]);
        v[6 + d] = wave_min(v[6 + d]);
        v[9 + d] = wave_max(v[9 + d]);
    }
    __shared__ float waves[4][12];
    if ((threadIdx.x & 63u) == 0)
        for (int k = 0; k < 12; k++) waves[threadIdx.x >> 6][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 12u) {
        const uint32_t k = threadIdx.x;
        const bool lower = k < 3u || (k >= 6u && k < 9u);
        float r = waves[0][k];
        for (int w = 1; w < 4; w++) r = lower ? mn(r, waves[w][k]) : mx(r, waves[w][k]);
        BigState &x = a.st[ch.big * kReplicas + blockIdx.x % kReplicas];
        if (lower) atomicMin(&x.box[k], enc(r));
        else atomicMax(&x.box[k], enc(r));
    }
}

__global__ __launch_bounds__(64) void big_merge_bounds_kernel(BigState *st) {
    const uint32_t k = threadIdx.x;
    if (k >= 12u) return;
    BigState *x = st + blockIdx.x * kReplicas;
    const bool lower = k < 3u || (k >= 6u && k < 9u);
    uint32_t r = x[0].box[k];
    for (uint32_t q = 1; q < kReplicas; q++) r = lower ? min(r, x[q].box[k]) : max(r, x[q].box[k]);
    x[0].box[k] = r;
}

__global__ __launch_bounds__(256) void big_bins_kernel(BigArgs a) {
    __shared__ uint32_t cnt[kSlots];
    __shared__ uint32_t bl[kSlots * 3], bh[kSlots * 3];
    const Chunk ch = a.chunks[blockIdx.x];
    const Bounds b = big_bounds(a.st[ch.big * kReplicas]);
    BigState &x = a.st[ch.big * kReplicas + blockIdx.x % kReplicas];
    const uint32_t t = threadIdx.x;
    for (uint32_t j = t; j < kSlots; j += 256u) cnt[j] = 0u;
    for (uint32_t j = t; j < kSlots * 3u; j += 256u) {
        bl[j] = enc(3.0e38f);
        bh[j] = enc(-3.0e38f);
    }
    __syncthreads();
    const uint32_t len = ch.end - ch.begin, run = (len + 255u) / 256u, r0 = ch.begin + min(len, t * run);
    bin_run(a.cur, r0, min(ch.end, r0 + run), a.blo, a.bhi, b, cnt, bl, bh);
    __syncthreads();
    for (uint32_t j = t; j < kSlots; j += 256u)
        if (cnt[j]) atomicAdd(&x.cnt[j], cnt[j]);
    for (uint32_t j = t; j < kSlots * 3u; j += 256u)
        if (cnt[j / 3u]) {
            atomicMin(&x.bl[j], bl[j]);
            atomicMax(&x.bh[j], bh[j]);
        }
}

// One wave per large node: the split (always: such a node holds more than 16
// triangles), its child pair, and large children listed for the next level.
__global__ __launch_bounds__(256) void big_split_kernel(BigArgs a, uint32_t big) {
    const BigNode bn = a.bigs[blockIdx.x];
    BigState *reps = a.st + blockIdx.x * kReplicas;
    BigState &x = reps[0];
    for (uint32_t k = threadIdx.x; k < kSlots * 7u; k += 256u) {
        if (k < kSlots) {
            uint32_t c = 0;
            for (uint32_t q = 0; q < kReplicas; q++) c += reps[q].cnt[k];
            x.cnt[k] = c;
        } else if (k < kSlots * 4u) {
            uint32_t j = k - kSlots, r = x.bl[j];
            for (uint32_t q = 1; q < kReplicas; q++) r = min(r, reps[q].bl[j]);
            x.bl[j] = r;
        } else {
            uint32_t j = k - kSlots * 4u, r = x.bh[j];
            for (uint32_t q = 1; q < kReplicas; q++) r = max(r, reps[q].bh[j]);
            x.bh[j] = r;
        }
    }
    __threadfence_block();
    __syncthreads();
    if (threadIdx.x >= 64u) return;
    const Bounds b = big_bounds(x);
    const uint32_t lane = threadIdx.x;
    int axis, bin;
    float cost;
    best_split(b, x.cnt, x.bl, x.bh, lane, axis, bin, cost);
    uint32_t L = 0;
    if (axis >= 0)
        for (int k = 0; k < bin; k++) L += x.cnt[axis * kBins + k];
    uint32_t left = (axis < 0 || L == 0 || L == bn.count) ? bn.count / 2u : L;
    if (lane == 0) {
        x.axis = axis;
        x.bin = bin;
        x.left = L;
        x.halve = axis < 0 ? 1u : 0u;
        uint32_t pair = atomicAdd(a.counter, 2u);
        a.nodes[bn.id] = lr_bvh_node{{b.lo[0], b.lo[1], b.lo[2]}, pair, {b.hi[0], b.hi[1], b.hi[2]}, 0u};
        const uint32_t f[2] = {bn.first, bn.first + left}, c[2] = {left, bn.count - left};
        for (int k = 0; k < 2; k++) {
            a.nodes[pair + k] = lr_bvh_node{{0, 0, 0}, f[k], {0, 0, 0}, c[k]};
            if (c[k] > big) {
                uint32_t at = atomicAdd(a.next_big_count, 1u);
                a.next_bigs[at] = BigNode{pair + (uint32_t)k, f[k], c[k], 0u, 0u};
            }
        }
    }
}

__device__ __forceinline__ bool goes_left(const BigArgs &a, uint32_t p, const BigState &x, const Bounds &b) {
    float lo = (&a.blo[p].x)[x.axis], hi = (&a.bhi[p].x)[x.axis];
    return bin_of(centroid(lo, hi), b.ml[x.axis], b.mh[x.axis]) < x.bin;
}

__global__ __launch_bounds__(256) void big_count_kernel(BigArgs a) {
    __shared__ uint32_t waves[4];
    const Chunk ch = a.chunks[blockIdx.x];
    const BigState &x = a.st[ch.big * kReplicas];
    const Bounds b = big_bounds(x);
    uint32_t sum = 0;
    for (uint32_t r0 = ch.begin; r0 < ch.end; r0 += 256u) {
        uint32_t i = r0 + threadIdx.x, total;
        bool left = !x.halve && i < ch.end && goes_left(a, a.cur[i], x, b);
        (void)block_prefix(left, waves, total);
        sum += total;
    }
    if (threadIdx.x == 0) a.chunk_left[blockIdx.x] = sum;
}

// Each large node's chunks: lefts before each chunk (one wave per node).
__global__ __launch_bounds__(64) void big_scan_kernel(BigArgs a) {
    const BigNode bn = a.bigs[blockIdx.x];
    const uint32_t lane = threadIdx.x;
    uint32_t base = 0;
    for (uint32_t c0 = 0; c0 < bn.chunks; c0 += 64u) {
        uint32_t c = c0 + lane;
        uint32_t v = c < bn.chunks ? a.chunk_left[bn.chunk0 + c] : 0u, inc = v;
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t t = __shfl_up(inc, o, 64);
            if ((int)lane >= o) inc += t;
        }
        if (c < bn.chunks) a.chunk_base[bn.chunk0 + c] = base + inc - v;
        base += __shfl(inc, 63, 64);
    }
}

template <bool SCATTER> __global__ __launch_bounds__(256) void big_partition_kernel(BigArgs a) {
    __shared__ uint32_t waves[4];
    const Chunk ch = a.chunks[blockIdx.x];
    const BigNode bn = a.bigs[ch.big];
    const BigState &x = a.st[ch.big * kReplicas];
    const Bounds b = big_bounds(x);
    if (x.halve) {
        if (SCATTER)
            for (uint32_t i = ch.begin + threadIdx.x; i < ch.end; i += 256u) a.next[i] = a.cur[i];
        return;
    }
    const uint32_t L = x.left, n = bn.count;
    uint32_t *tab = a.tab + bn.first;
    uint32_t base = a.chunk_base[blockIdx.x];
    for (uint32_t r0 = ch.begin; r0 < ch.end; r0 += 256u) {
        uint32_t i = r0 + threadIdx.x, total;
        uint32_t p = i < ch.end ? a.cur[i] : 0u;
        bool left = i < ch.end && goes_left(a, p, x, b);
        uint32_t lb = base + block_prefix(left, waves, total);
        if (i < ch.end) {
            uint32_t at = i - bn.first;
            if (SCATTER) {
                uint32_t d = destination(left, at, lb, L, n, tab);
                if (d < n) a.next[bn.first + d] = p;
            } else {
                table_entry(left, at, lb, L, n, tab);
            }
        }
        base += total;
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
    DeviceArray<uint32_t> ord[2], tab, counter, inner, rank, index, chunk_left, chunk_base, big_count;
    DeviceArray<lr_bvh_node> tmp;
    DeviceArray<BigNode> bigs_d, next_bigs_d;
    DeviceArray<Chunk> chunks_d;
    DeviceArray<BigState> states;
    int rc;
    const size_t cap = 2 * n;
    // Large nodes of one level are disjoint: at most n / kBigNode of them.
    const size_t max_big = n / kBigNode + 1, max_chunks = n / kChunk + max_big + 1;
    if ((rc = blo.alloc(n, "hipMalloc(build boxes)")) || (rc = bhi.alloc(n, "hipMalloc(build boxes)")) ||
        (rc = ord[0].alloc(n, "hipMalloc(build order)")) || (rc = ord[1].alloc(n, "hipMalloc(build order)")) ||
        (rc = tab.alloc(n, "hipMalloc(build tables)")) || (rc = counter.alloc(1, "hipMalloc(build counter)")) ||
        (rc = tmp.alloc(cap, "hipMalloc(build nodes)")) || (rc = bigs_d.alloc(max_big, "hipMalloc(build)")) ||
        (rc = next_bigs_d.alloc(max_big, "hipMalloc(build)")) ||
        (rc = states.alloc(max_big * kReplicas, "hipMalloc(build)")) ||
        (rc = chunks_d.alloc(max_chunks, "hipMalloc(build)")) || (rc = chunk_left.alloc(max_chunks, "hipMalloc(build)")) ||
        (rc = chunk_base.alloc(max_chunks, "hipMalloc(build)")) || (rc = big_count.alloc(1, "hipMalloc(build)")))
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
    std::vector<BigNode> bigs;
    std::vector<Chunk> chunks;
    if (n > kBigNode) bigs.push_back(BigNode{0u, 0u, n32, 0u, 0u});
    int c = 0;
    for (;;) {
        auto [begin, end] = levels.back();
        const uint32_t nbig = (uint32_t)bigs.size();
        if (nbig) {
            chunks.clear();
            for (uint32_t j = 0; j < nbig; j++) {
                BigNode &bn = bigs[j];
                bn.chunk0 = (uint32_t)chunks.size();
                bn.chunks = (bn.count + kChunk - 1) / kChunk;
                for (uint32_t q = 0; q < bn.chunks; q++)
                    chunks.push_back(Chunk{j, bn.first + q * kChunk, bn.first + std::min(bn.count, (q + 1) * kChunk)});
            }
            if (nbig > max_big || chunks.size() > max_chunks)
                return fail(LR_DEVICE_ERROR, "BVH build: more large nodes than a level can hold");
            const uint32_t nch = (uint32_t)chunks.size();
            if ((e = hipMemcpyAsync(bigs_d.p, bigs.data(), nbig * sizeof(BigNode), hipMemcpyHostToDevice, s)) != hipSuccess ||
                (e = hipMemcpyAsync(chunks_d.p, chunks.data(), nch * sizeof(Chunk), hipMemcpyHostToDevice, s)) != hipSuccess ||
                (e = hipMemsetAsync(big_count.p, 0, 4, s)) != hipSuccess)
                return hip_fail(e, "hipMemcpy(build large nodes)");
            BigArgs ba{blo.p, bhi.p, ord[c].p, ord[c ^ 1].p, tab.p, tmp.p, counter.p, bigs_d.p, chunks_d.p, states.p,
                       chunk_left.p, chunk_base.p, next_bigs_d.p, big_count.p};
            hipLaunchKernelGGL(big_init_kernel, dim3(nbig * kReplicas), dim3(64), 0, s, states.p);
            hipLaunchKernelGGL(big_bounds_kernel, dim3(nch), dim3(256), 0, s, ba);
            hipLaunchKernelGGL(big_merge_bounds_kernel, dim3(nbig), dim3(64), 0, s, states.p);
            hipLaunchKernelGGL(big_bins_kernel, dim3(nch), dim3(256), 0, s, ba);
            hipLaunchKernelGGL(big_split_kernel, dim3(nbig), dim3(256), 0, s, ba, kBigNode);
            hipLaunchKernelGGL(big_count_kernel, dim3(nch), dim3(256), 0, s, ba);
            hipLaunchKernelGGL(big_scan_kernel, dim3(nbig), dim3(64), 0, s, ba);
            hipLaunchKernelGGL(big_partition_kernel<false>, dim3(nch), dim3(256), 0, s, ba);
            hipLaunchKernelGGL(big_partition_kernel<true>, dim3(nch), dim3(256), 0, s, ba);
        }
        // One node per wave while that fills the GPU, up to 64 per wave beyond.
        const uint32_t batch = std::min<uint32_t>(64u, std::max<uint32_t>(1u, (end - begin) / 16384u));
        LevelArgs a{blo.p, bhi.p, ord[c].p, ord[c ^ 1].p, order, tab.p, tmp.p, counter.p, begin, end, kBigNode, batch};
        uint32_t blocks = std::min<uint32_t>((end - begin + batch - 1) / batch, 65536u);
        hipLaunchKernelGGL(level_kernel, dim3(blocks), dim3(64), 0, s, a);
        uint32_t used = 0, next_big = 0;
        if ((e = hipMemcpyAsync(&used, counter.p, 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (nbig && (e = hipMemcpyAsync(&next_big, big_count.p, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            return hip_fail(e, "BVH build level");
        if (used > cap) return fail(LR_DEVICE_ERROR, "BVH build handed out more nodes than the tree can hold");
        if (next_big > max_big) return fail(LR_DEVICE_ERROR, "BVH build: more large nodes than a level can hold");
        bigs.resize(next_big);
        if (next_big && ((e = hipMemcpyAsync(bigs.data(), next_bigs_d.p, next_big * sizeof(BigNode), hipMemcpyDeviceToHost,
                                             s)) != hipSuccess ||
                         (e = hipStreamSynchronize(s)) != hipSuccess))
            return hip_fail(e, "hipMemcpy(build large nodes)");
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
