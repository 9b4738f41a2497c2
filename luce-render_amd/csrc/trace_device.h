// Device-side BVH traversal over the wide layouts (bvh_wide.cpp), shared by
// the probe, ray-batch, shadow and integrator kernels.
//
// Semantics follow shaders/bvh.glsl:
//   * the slab test box_entry (bvh.glsl:17-25) with the entry clamped at 0 and
//     the exit at the current best t, a child taken when enter <= leave;
//   * Moller-Trumbore triangle_hit (bvh.glsl:27-47): |det| < 1e-20 rejected,
//     u outside [0,1], v < 0 or u + v > 1 rejected, t <= 0 or t >= best
//     rejected (the first triangle found at a given t keeps it);
//   * axis-parallel directions replaced by +-1e-20 before inverting
//     (bvh.glsl:55-56);
//   * closest hit: hit children visited nearest first (ties: lower child
//     index), the others pushed farthest first (bvh.glsl:73-84); with width 2
//     this is the reference's visiting order exactly; wider nodes visit the
//     same boxes in another order, so where two triangles tie at exactly the
//     same t they settle the tie as the binary order would
//     (binary_first);
//   * any hit: children taken in child order, the first triangle hit ends it
//     (bvh.glsl:93-122).
// Unlike bvh.glsl:79 no stack entry is ever dropped: entries past the LDS
// short stack spill to a global area sized by the tree's exact stack bound.
//
// All arithmetic is f32 in the reference's operation order; the library is
// compiled with -ffp-contract=off so no multiply-add is fused, and divisions
// and square roots are IEEE correctly rounded (HIP's default).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lr {

struct DeviceScene {
    const uint8_t *nodes;    // wide nodes, node_bytes each
    const float *tris;       // 9 floats per leaf slot (36 B): p0.xyz e1.xyz e2.xyz
    const uint32_t *faces;   // polygon per leaf slot
    float root_lo[3];
    float root_hi[3];
    uint32_t empty;          // an empty tree: every ray misses
    uint32_t *spill;         // overflow stack entries, [depth - S][thread]
    uint32_t spill_stride;   // threads the launch covers
    uint32_t spill_depth;    // entries per thread the spill area holds
    uint32_t *work;          // persistent launches: {rays handed out, waves finished}, zero between launches
    const uint8_t *binary;   // the width-2 nodes (the binary tree), for wider traversals' ties
    const uint2 *ranges;     // their leaf-order triangle slot ranges
};

struct Hit {
    float t;
    uint32_t slot;  // leaf-order triangle, 0xffffffff when missed
    float u, v;
};

constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kLeafFlag = 0x80000000u;

__device__ __forceinline__ uint32_t leaf_entry(uint32_t first, uint32_t count) {
    return kLeafFlag | ((count - 1u) << 27) | first;
}

__device__ __forceinline__ float safe_inverse(float d) {
    float s = d;
    if (fabsf(d) < 1e-20f) {
        float sg = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
        s = sg * 1e-20f + (d == 0.0f ? 1.0f : 0.0f) * 1e-20f;
    }
    return 1.0f / s;
}

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    float s = ax * bx;
    s = s + ay * by;
    return s + az * bz;
}

// box_entry (bvh.glsl:17-25); -1 when missed.
__device__ __forceinline__ float box_entry(float lox, float loy, float loz, float hix, float hiy, float hiz,
                                           float ox, float oy, float oz, float ix, float iy, float iz, float tmax) {
    float ax = (lox - ox) * ix, ay = (loy - oy) * iy, az = (loz - oz) * iz;
    float bx = (hix - ox) * ix, by = (hiy - oy) * iy, bz = (hiz - oz) * iz;
    float enter = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), 0.0f));
    float leave = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tmax));
    return enter <= leave ? enter : -1.0f;
}

// A leaf slot's repacked triangle as three registers of four (the last holds
// e2.z alone): 36 bytes a slot, read at 4-byte alignment.
constexpr uint32_t kTriangleFloats = 9;
__device__ __forceinline__ void load_triangle(const float *tris, uint32_t slot, float4 &a, float4 &b, float4 &c) {
    const float *p = tris + (size_t)slot * kTriangleFloats;
    a = make_float4(p[0], p[1], p[2], p[3]);
    b = make_float4(p[4], p[5], p[6], p[7]);
    c = make_float4(p[8], 0.0f, 0.0f, 0.0f);
}

// triangle_hit (bvh.glsl:27-47) on a repacked triangle: whether it is hit
// before `best`, and where; `tie_out`: hit exactly at `best`. Every quantity is
// computed exactly as the reference computes it; its early exits become one
// predicate over the same comparisons, so the loads and arithmetic run
// straight through (no divergent branch, no load held back behind one).
__device__ __forceinline__ bool triangle_test(float4 a, float4 b, float4 c, float ox, float oy, float oz, float dx,
                                              float dy, float dz, float best, float &t_out, float &u_out, float &v_out,
                                              bool &tie_out) {
    float e1x = a.w, e1y = b.x, e1z = b.y;
    float e2x = b.z, e2y = b.w, e2z = c.x;
    float px = dy * e2z - dz * e2y, py = dz * e2x - dx * e2z, pz = dx * e2y - dy * e2x;
    float det = dot3(e1x, e1y, e1z, px, py, pz);
    float inv = 1.0f / det;
    float sx = ox - a.x, sy = oy - a.y, sz = oz - a.z;
    float u = dot3(sx, sy, sz, px, py, pz) * inv;
    float qx = sy * e1z - sz * e1y, qy = sz * e1x - sx * e1z, qz = sx * e1y - sy * e1x;
    float v = dot3(dx, dy, dz, qx, qy, qz) * inv;
    float t = dot3(e2x, e2y, e2z, qx, qy, qz) * inv;
    bool reject = fabsf(det) < 1e-20f;
    reject = reject || u < 0.0f || u > 1.0f;
    reject = reject || v < 0.0f || u + v > 1.0f;
    reject = reject || t <= 0.0f;
    tie_out = !reject && t == best;
    reject = reject || t >= best;
    t_out = t;
    u_out = u;
    v_out = v;
    return !reject;
}

// Layout codes: 2, 4, 8 the float layouts of those widths; kQuantizedLayout
// + 4 and + 8 the 8-bit quantized layouts (a node's child boxes as bytes on a
// per-axis power-of-two grid from the node's own low corner, rounded outward:
// 64 and 128 bytes a node instead of 128 and 256).
constexpr int kQuantizedLayout = 16;
template <int W> constexpr int kWidthOf = W & 15;
template <int W> constexpr bool kQuantizedOf = (W & kQuantizedLayout) != 0;

template <int W> struct NodeView {
    static constexpr int kQuads = (W == 2) ? 4 : (W == 4 ? 8 : 16);
    float4 raw[kQuads];
    __device__ __forceinline__ void load(const uint8_t *nodes, uint32_t index) {
        const float4 *p = reinterpret_cast<const float4 *>(nodes + (size_t)index * (kQuads * 16));
#pragma unroll
        for (int q = 0; q < kQuads; q++) raw[q] = p[q];
    }
    __device__ __forceinline__ float f(int i) const {
        const float *x = reinterpret_cast<const float *>(raw);
        return x[i];
    }
    __device__ __forceinline__ uint32_t u(int i) const {
        const uint32_t *x = reinterpret_cast<const uint32_t *>(raw);
        return x[i];
    }
    __device__ __forceinline__ float lo_x(int c) const { return f(c); }
    __device__ __forceinline__ float hi_x(int c) const { return f(W + c); }
    __device__ __forceinline__ float lo_y(int c) const { return f(2 * W + c); }
    __device__ __forceinline__ float hi_y(int c) const { return f(3 * W + c); }
    __device__ __forceinline__ float lo_z(int c) const { return f(4 * W + c); }
    __device__ __forceinline__ float hi_z(int c) const { return f(5 * W + c); }
    __device__ __forceinline__ uint32_t child(int c) const { return u(6 * W + c); }
    __device__ __forceinline__ uint32_t count(int c) const { return (u(7 * W + c / 4) >> (8 * (c % 4))) & 0xffu; }
};

// A quantized node (width WD): origin xyz and three biased exponent bytes
// (16 B), then the six planes' bytes per child (lo x, hi x, lo y, hi y, lo z,
// hi z; WD each), the WD child words and WD count bytes. A plane decodes to
// origin + q * 2^(e - 127), computed in f32 exactly as the quantizer checked
// it, so every decoded box holds the child's float box.
template <int WD> struct QuantNodeView {
    static constexpr int kQuads = WD == 4 ? 4 : 8;
    float4 raw[kQuads];
    float ox, oy, oz, sx, sy, sz;
    __device__ __forceinline__ void load(const uint8_t *nodes, uint32_t index) {
        const float4 *p = reinterpret_cast<const float4 *>(nodes + (size_t)index * (kQuads * 16));
#pragma unroll
        for (int q = 0; q < kQuads; q++) raw[q] = p[q];
        ox = raw[0].x, oy = raw[0].y, oz = raw[0].z;
        uint32_t e = __float_as_uint(raw[0].w);
        sx = __uint_as_float((e & 0xffu) << 23);
        sy = __uint_as_float(((e >> 8) & 0xffu) << 23);
        sz = __uint_as_float(((e >> 16) & 0xffu) << 23);
    }
    __device__ __forceinline__ uint32_t u(int i) const {
        const uint32_t *x = reinterpret_cast<const uint32_t *>(raw);
        return x[i];
    }
    __device__ __forceinline__ float byte(int i) const { return (float)((u(4 + i / 4) >> (8 * (i % 4))) & 0xffu); }
    __device__ __forceinline__ float lo_x(int c) const { return ox + byte(c) * sx; }
    __device__ __forceinline__ float hi_x(int c) const { return ox + byte(WD + c) * sx; }
    __device__ __forceinline__ float lo_y(int c) const { return oy + byte(2 * WD + c) * sy; }
    __device__ __forceinline__ float hi_y(int c) const { return oy + byte(3 * WD + c) * sy; }
    __device__ __forceinline__ float lo_z(int c) const { return oz + byte(4 * WD + c) * sz; }
    __device__ __forceinline__ float hi_z(int c) const { return oz + byte(5 * WD + c) * sz; }
    __device__ __forceinline__ uint32_t child(int c) const { return u(4 + 6 * WD / 4 + c); }
    __device__ __forceinline__ uint32_t count(int c) const {
        return (u(4 + 6 * WD / 4 + WD + c / 4) >> (8 * (c % 4))) & 0xffu;
    }
};

template <int W, bool Q = kQuantizedOf<W>> struct NodeOfT {
    using type = NodeView<W>;
};
template <int W> struct NodeOfT<W, true> {
    using type = QuantNodeView<kWidthOf<W>>;
};
template <int W> using NodeOf = typename NodeOfT<W>::type;

// The per-lane stack: S entries in LDS (lane-interleaved, `stride` apart),
// the rest in the global spill area.
template <int S> struct LaneStack {
    uint32_t *lds;
    uint32_t *spill;
    uint32_t spill_stride;
    int sp;
    __device__ __forceinline__ void push(uint32_t e, int stride) {
        if (sp < S) lds[sp * stride] = e;
        else spill[(size_t)(sp - S) * spill_stride] = e;
        sp++;
    }
    __device__ __forceinline__ uint32_t pop(int stride) {
        sp--;
        return sp < S ? lds[sp * stride] : spill[(size_t)(sp - S) * spill_stride];
    }
};

// Of two leaf-order triangles the ray hits at the same t, whether `a` is the
// one the reference's binary traversal finds first (and so keeps): below
// their lowest common node it visits the child with the nearer box entry
// first (ties: the left one) and finishes it before the other; within one
// leaf it tests triangles in order.
__device__ __forceinline__ bool binary_first(const DeviceScene &sc, uint32_t a, uint32_t b, float ox, float oy, float oz,
                                          float ix, float iy, float iz, float tmax) {
    uint32_t node = 0;
    for (uint32_t depth = 0; depth < 4096u; depth++) {
        NodeView<2> n;
        n.load(sc.binary, node);
        uint32_t lo[2], hi[2];
        for (int c = 0; c < 2; c++) {
            uint32_t ch = n.child(c), cnt = n.count(c);
            if (cnt) {
                lo[c] = ch;
                hi[c] = ch + cnt - 1u;
            } else if (ch != kNone) {
                uint2 r = sc.ranges[ch];
                lo[c] = r.x;
                hi[c] = r.y;
            } else {
                lo[c] = 1u;
                hi[c] = 0u;
            }
        }
        int ca = (a >= lo[0] && a <= hi[0]) ? 0 : 1, cb = (b >= lo[0] && b <= hi[0]) ? 0 : 1;
        if (ca != cb) {
            float e0 = box_entry(n.lo_x(0), n.lo_y(0), n.lo_z(0), n.hi_x(0), n.hi_y(0), n.hi_z(0), ox, oy, oz, ix, iy,
                                 iz, tmax);
            float e1 = box_entry(n.lo_x(1), n.lo_y(1), n.lo_z(1), n.hi_x(1), n.hi_y(1), n.hi_z(1), ox, oy, oz, ix, iy,
                                 iz, tmax);
            bool left_first = e0 >= 0.0f && (e1 < 0.0f || e0 <= e1);
            return (ca == 0) == left_first;
        }
        if (n.count(ca)) return a < b;  // the same leaf
        node = n.child(ca);
    }
    return a < b;
}

// One ray's closest-hit walk (bvh.glsl:51-90) over width-W nodes, advanced one
// node at a time, so a persistent wave can hand a finished lane a new ray.
template <int W, int S> struct ClosestWalk {
    static constexpr int WD = kWidthOf<W>;
    float ox, oy, oz, dx, dy, dz, ix, iy, iz;
    Hit hit;
    uint32_t cur;
    LaneStack<S> stack;

    // Sets the ray up; false when it misses the root box (or the tree is
    // empty), and hit then holds the miss.
    __device__ __forceinline__ bool begin(const DeviceScene &sc, float ox_, float oy_, float oz_, float dx_, float dy_,
                                          float dz_, float tmax) {
        ox = ox_, oy = oy_, oz = oz_, dx = dx_, dy = dy_, dz = dz_;
        hit = Hit{tmax, kNone, 0.0f, 0.0f};
        ix = safe_inverse(dx), iy = safe_inverse(dy), iz = safe_inverse(dz);
        cur = 0;
        stack.sp = 0;
        return !(sc.empty || sc.root_lo[0] > sc.root_hi[0] ||
                 box_entry(sc.root_lo[0], sc.root_lo[1], sc.root_lo[2], sc.root_hi[0], sc.root_hi[1], sc.root_hi[2],
                           ox, oy, oz, ix, iy, iz, hit.t) < 0.0f);
    }

    // Visits the current node: a leaf's triangles, or an interior node's
    // children (hit ones nearest first, ties by child index, the others
    // pushed farthest first). False once the stack is empty.
    __device__ __forceinline__ bool step(const DeviceScene &sc, int stride) {
        if (cur & kLeafFlag) {
            // The leaf's triangles in order; the next one's loads are issued
            // before the current one is tested.
            uint32_t first = cur & 0x07ffffffu, last = first + ((cur >> 27) & 0xfu);
            float4 a, b, c;
            load_triangle(sc.tris, first, a, b, c);
            for (uint32_t slot = first;; slot++) {
                float4 na, nb, nc;
                load_triangle(sc.tris, slot < last ? slot + 1 : slot, na, nb, nc);
                float t, u, v;
                bool tie;
                bool take = triangle_test(a, b, c, ox, oy, oz, dx, dy, dz, hit.t, t, u, v, tie);
                if (WD > 2 && tie && hit.slot != kNone)
                    take = binary_first(sc, slot, hit.slot, ox, oy, oz, ix, iy, iz, hit.t);
                if (take) hit = Hit{t, slot, u, v};
                if (slot == last) break;
                a = na, b = nb, c = nc;
            }
        } else {
            NodeOf<W> n;
            n.load(sc.nodes, cur);
            float dist[WD];
            uint32_t ent[WD];
            int key[WD];
            int hits = 0;
#pragma unroll
            for (int c = 0; c < WD; c++) {
                uint32_t ch = n.child(c);
                uint32_t cnt = n.count(c);
                // every slot's planes are tested (an empty slot holds zeros)
                // and empty slots are masked afterwards
                float e = box_entry(n.lo_x(c), n.lo_y(c), n.lo_z(c), n.hi_x(c), n.hi_y(c), n.hi_z(c), ox, oy, oz, ix,
                                    iy, iz, hit.t);
                bool take = ch != kNone && e >= 0.0f;
                hits += take ? 1 : 0;
                dist[c] = take ? e : __builtin_huge_valf();
                ent[c] = cnt ? leaf_entry(ch, cnt) : ch;
                key[c] = c;
            }
            if (hits > 0) {
                // Sort by (entry distance, child index): a stable order.
#pragma unroll
                for (int i = 1; i < WD; i++) {
#pragma unroll
                    for (int j = WD - 1; j >= i; j--) {
                        bool sw = dist[j] < dist[j - 1] || (dist[j] == dist[j - 1] && key[j] < key[j - 1]);
                        float td = dist[j];
                        uint32_t te = ent[j];
                        int tk = key[j];
                        dist[j] = sw ? dist[j - 1] : dist[j];
                        ent[j] = sw ? ent[j - 1] : ent[j];
                        key[j] = sw ? key[j - 1] : key[j];
                        dist[j - 1] = sw ? td : dist[j - 1];
                        ent[j - 1] = sw ? te : ent[j - 1];
                        key[j - 1] = sw ? tk : key[j - 1];
                    }
                }
#pragma unroll
                for (int i = WD - 1; i >= 1; i--)
                    if (i < hits) stack.push(ent[i], stride);
                cur = ent[0];
                return true;
            }
        }
        if (stack.sp == 0) return false;
        cur = stack.pop(stride);
        return true;
    }
};

// One ray's any-hit walk (bvh.glsl:93-122), advanced one node at a time.
template <int W, int S> struct AnyWalk {
    static constexpr int WD = kWidthOf<W>;
    float ox, oy, oz, dx, dy, dz, ix, iy, iz, tmax;
    uint32_t cur;
    bool blocked;
    LaneStack<S> stack;

    __device__ __forceinline__ bool begin(const DeviceScene &sc, float ox_, float oy_, float oz_, float dx_, float dy_,
                                          float dz_, float tmax_) {
        ox = ox_, oy = oy_, oz = oz_, dx = dx_, dy = dy_, dz = dz_, tmax = tmax_;
        ix = safe_inverse(dx), iy = safe_inverse(dy), iz = safe_inverse(dz);
        cur = 0;
        blocked = false;
        stack.sp = 0;
        return !(sc.empty || sc.root_lo[0] > sc.root_hi[0] ||
                 box_entry(sc.root_lo[0], sc.root_lo[1], sc.root_lo[2], sc.root_hi[0], sc.root_hi[1], sc.root_hi[2],
                           ox, oy, oz, ix, iy, iz, tmax) < 0.0f);
    }

    // Children in child order: the first taken is visited, the others pushed
    // so that they pop in child order. False once blocked or exhausted.
    __device__ __forceinline__ bool step(const DeviceScene &sc, int stride) {
        if (cur & kLeafFlag) {
            uint32_t first = cur & 0x07ffffffu, last = first + ((cur >> 27) & 0xfu);
            float4 a, b, c;
            load_triangle(sc.tris, first, a, b, c);
            for (uint32_t slot = first;; slot++) {
                float4 na, nb, nc;
                load_triangle(sc.tris, slot < last ? slot + 1 : slot, na, nb, nc);
                float t, u, v;
                bool tie;
                if (triangle_test(a, b, c, ox, oy, oz, dx, dy, dz, tmax, t, u, v, tie)) {
                    blocked = true;
                    return false;
                }
                if (slot == last) break;
                a = na, b = nb, c = nc;
            }
        } else {
            NodeOf<W> n;
            n.load(sc.nodes, cur);
            uint32_t ent[WD];
            bool take[WD];
            int hits = 0;
#pragma unroll
            for (int c = 0; c < WD; c++) {
                uint32_t ch = n.child(c);
                uint32_t cnt = n.count(c);
                float e = box_entry(n.lo_x(c), n.lo_y(c), n.lo_z(c), n.hi_x(c), n.hi_y(c), n.hi_z(c), ox, oy, oz, ix,
                                    iy, iz, tmax);
                take[c] = ch != kNone && e >= 0.0f;
                hits += take[c] ? 1 : 0;
                ent[c] = cnt ? leaf_entry(ch, cnt) : ch;
            }
            if (hits > 0) {
                uint32_t next = kNone;
#pragma unroll
                for (int c = WD - 1; c >= 0; c--) {
                    if (!take[c]) continue;
                    if (next != kNone) stack.push(next, stride);
                    next = ent[c];
                }
                cur = next;
                return true;
            }
        }
        if (stack.sp == 0) return false;
        cur = stack.pop(stride);
        return true;
    }
};

// trace_closest (bvh.glsl:51-90): the walk run to its end.
template <int W, int S>
__device__ __forceinline__ Hit trace_closest(const DeviceScene &sc, float ox, float oy, float oz, float dx, float dy,
                                             float dz, float tmax, uint32_t *lds, int stride, uint32_t thread) {
    ClosestWalk<W, S> walk;
    walk.stack = LaneStack<S>{lds, sc.spill + thread, sc.spill_stride, 0};
    if (walk.begin(sc, ox, oy, oz, dx, dy, dz, tmax))
        while (walk.step(sc, stride)) {
        }
    return walk.hit;
}

// trace_any (bvh.glsl:93-122): whether anything blocks the ray before tmax.
template <int W, int S>
__device__ __forceinline__ bool trace_any(const DeviceScene &sc, float ox, float oy, float oz, float dx, float dy,
                                          float dz, float tmax, uint32_t *lds, int stride, uint32_t thread) {
    AnyWalk<W, S> walk;
    walk.stack = LaneStack<S>{lds, sc.spill + thread, sc.spill_stride, 0};
    if (walk.begin(sc, ox, oy, oz, dx, dy, dz, tmax))
        while (walk.step(sc, stride)) {
        }
    return walk.blocked;
}

// Persistent launches: every wave takes rays from a shared counter and, once
// at least `refill` of its lanes have finished their rays, hands those lanes
// the next rays, so lanes do not idle behind the wave's slowest ray. Each
// ray's walk is the one-ray-per-thread walk, node for node. The source's
// begin(sc, walk, r) sets ray r up (false: it is already done and stored) and
// finish(sc, walk, r) stores its result. Finishing waits for the refill, so
// the lanes whose rays ended finish together rather than one or two at a
// time. sc.work holds the counter pair, zero between launches: the last wave
// of a launch (gridDim.x workgroups of G threads) clears it.
template <int G, class Walk, class Source>
__device__ __forceinline__ void stream_rays(const DeviceScene &sc, Walk &walk, Source &src, uint32_t n,
                                            uint32_t refill) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t ray = kNone;
    bool walking = false;
    bool more = true;
    for (;;) {
        uint64_t idle = __ballot(!walking);
        uint32_t count = (uint32_t)__popcll(idle);
        if (count >= refill || (!more && count == 64u)) {
            if (ray != kNone && !walking) {
                src.finish(sc, walk, ray);
                ray = kNone;
            }
            if (!more && count == 64u) break;
            if (more) {
                uint32_t leader = (uint32_t)__ffsll((unsigned long long)idle) - 1u;
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(&sc.work[0], count);
                base = __shfl(base, (int)leader);
                more = (uint64_t)base + count < n;
                if (!walking) {
                    uint32_t r = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                    if (r < n && src.begin(sc, walk, r)) {
                        ray = r;
                        walking = true;
                    }
                }
            }
        }
        if (walking) walking = walk.step(sc, G);
    }
    // The last wave out clears the counters for the next launch.
    if (lane == 0u) {
        uint32_t waves = gridDim.x * (G / 64u);
        if (atomicAdd(&sc.work[1], 1u) == waves - 1u) {
            atomicExch(&sc.work[0], 0u);
            atomicExch(&sc.work[1], 0u);
        }
    }
}

}  // namespace lr
