// Tracing kernels: the primary-ray probe (shaders/probe.comp), closest hits of
// ray batches and any-hit shadow tests, each instantiated per BVH width W,
// LDS stack depth S and workgroup size G so the autotuner can pick a variant.
#include "trace_kernels.h"

#include <map>
#include <mutex>
#include <utility>

#include "lr_scene.h"

// Kernels receive the scene by value; each launcher stamps the spill stride
// (the threads it launches) before launching.

namespace lr {

struct Frame {
    float v[16];
};

// probe.comp:26-30: the pixel-centre ray, f32 in the reference's order.
__device__ __forceinline__ void probe_ray(const Frame &f, uint32_t x, uint32_t y, float &ox, float &oy, float &oz,
                                          float &dx, float &dy, float &dz) {
    float fx = (float)x + 0.5f, fy = (float)y + 0.5f;
    float filmx = f.v[4] + fx * f.v[8];
    filmx = filmx + fy * f.v[12];
    float filmy = f.v[5] + fx * f.v[9];
    filmy = filmy + fy * f.v[13];
    float filmz = f.v[6] + fx * f.v[10];
    filmz = filmz + fy * f.v[14];
    if (f.v[3] > 0.5f) {
        ox = f.v[0] + filmx;
        oy = f.v[1] + filmy;
        oz = f.v[2] + filmz;
        float cx = f.v[9] * f.v[14] - f.v[10] * f.v[13];
        float cy = f.v[10] * f.v[12] - f.v[8] * f.v[14];
        float cz = f.v[8] * f.v[13] - f.v[9] * f.v[12];
        float l = sqrtf(dot3(cx, cy, cz, cx, cy, cz));
        dx = cx / l;
        dy = cy / l;
        dz = cz / l;
    } else {
        ox = f.v[0];
        oy = f.v[1];
        oz = f.v[2];
        dx = filmx;
        dy = filmy;
        dz = filmz;
    }
}

// Workgroups go to the 8 XCDs round-robin; renumber them so neighbouring
// tiles (which share BVH nodes) land on the same XCD and its L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t block, uint32_t blocks) {
    uint32_t per = blocks / 8u;
    if (per == 0u || block >= per * 8u) return block;
    return (block % 8u) * per + block / 8u;
}

template <int W, int S, int G>
__global__ __launch_bounds__(G) void probe_kernel(DeviceScene sc, Frame f, uint32_t width, uint32_t height,
                                                  uint2 *out) {
    __shared__ uint32_t lds[S * G];
    uint32_t block = xcd_block(blockIdx.x, gridDim.x);
    uint32_t gid = block * G + threadIdx.x;
    // 8x8 pixel tiles per 64 lanes, tiles row-major.
    uint32_t tile = gid >> 6, lane = gid & 63u;
    uint32_t tiles_x = (width + 7u) / 8u;
    uint32_t x = (tile % tiles_x) * 8u + (lane & 7u), y = (tile / tiles_x) * 8u + (lane >> 3);
    if (x >= width || y >= height) return;
    float ox, oy, oz, dx, dy, dz;
    probe_ray(f, x, y, ox, oy, oz, dx, dy, dz);
    Hit h = trace_closest<W, S>(sc, ox, oy, oz, dx, dy, dz, 3.0e38f, lds + threadIdx.x, G, blockIdx.x * G + threadIdx.x);
    uint32_t face = h.slot == kNone ? kNone : sc.faces[h.slot];
    out[(size_t)y * width + x] = make_uint2(face, __float_as_uint(h.t));
}

template <int W, int S, int G>
__global__ __launch_bounds__(G) void closest_kernel(DeviceScene sc, const float4 *rays, uint32_t n, uint2 *out) {
    __shared__ uint32_t lds[S * G];
    uint32_t block = xcd_block(blockIdx.x, gridDim.x);
    uint32_t i = block * G + threadIdx.x;
    if (i >= n) return;
    float4 a = rays[(size_t)i * 2], b = rays[(size_t)i * 2 + 1];
    Hit h = trace_closest<W, S>(sc, a.x, a.y, a.z, b.x, b.y, b.z, a.w, lds + threadIdx.x, G, blockIdx.x * G + threadIdx.x);
    uint32_t face = h.slot == kNone ? kNone : sc.faces[h.slot];
    out[i] = make_uint2(face, __float_as_uint(h.t));
}

template <int W, int S, int G>
__global__ __launch_bounds__(G) void any_kernel(DeviceScene sc, const float4 *rays, uint32_t n, uint32_t *out) {
    __shared__ uint32_t lds[S * G];
    uint32_t block = xcd_block(blockIdx.x, gridDim.x);
    uint32_t i = block * G + threadIdx.x;
    if (i >= n) return;
    float4 a = rays[(size_t)i * 2], b = rays[(size_t)i * 2 + 1];
    out[i] = trace_any<W, S>(sc, a.x, a.y, a.z, b.x, b.y, b.z, a.w, lds + threadIdx.x, G, blockIdx.x * G + threadIdx.x)
                 ? 1u
                 : 0u;
}

// ---------------------------------------------------------------- persistent

// Rays in the batch layout: (origin, t_max), (direction, -).
struct BatchClosest {
    const float4 *rays;
    uint2 *out;
    template <class Walk> __device__ __forceinline__ bool begin(const DeviceScene &sc, Walk &w, uint32_t r) {
        float4 a = rays[(size_t)r * 2], b = rays[(size_t)r * 2 + 1];
        if (w.begin(sc, a.x, a.y, a.z, b.x, b.y, b.z, a.w)) return true;
        out[r] = make_uint2(kNone, __float_as_uint(a.w));
        return false;
    }
    template <class Walk> __device__ __forceinline__ void finish(const DeviceScene &sc, Walk &w, uint32_t r) {
        uint32_t face = w.hit.slot == kNone ? kNone : sc.faces[w.hit.slot];
        out[r] = make_uint2(face, __float_as_uint(w.hit.t));
    }
};

struct BatchAny {
    const float4 *rays;
    uint32_t *out;
    template <class Walk> __device__ __forceinline__ bool begin(const DeviceScene &sc, Walk &w, uint32_t r) {
        float4 a = rays[(size_t)r * 2], b = rays[(size_t)r * 2 + 1];
        if (w.begin(sc, a.x, a.y, a.z, b.x, b.y, b.z, a.w)) return true;
        out[r] = 0u;
        return false;
    }
    template <class Walk> __device__ __forceinline__ void finish(const DeviceScene &, Walk &w, uint32_t r) {
        out[r] = w.blocked ? 1u : 0u;
    }
};

template <int W, int S, int G>
__global__ __launch_bounds__(G) void closest_stream_kernel(DeviceScene sc, const float4 *rays, uint32_t n, uint2 *out,
                                                           uint32_t refill) {
    __shared__ uint32_t lds[S * G];
    uint32_t thread = blockIdx.x * G + threadIdx.x;
    ClosestWalk<W, S> walk;
    walk.stack = LaneStack<S>{lds + threadIdx.x, sc.spill + thread, sc.spill_stride, 0};
    BatchClosest src{rays, out};
    stream_rays<G>(sc, walk, src, n, refill);
}

// A frame's primary rays, then a batch of rays, as one stream of items: the
// first `primary` items (probe_items(width, height)) are the frame's 8x8
// pixel tiles, 64 items each, in the probe kernel's order (items past the
// image edge end at once, storing nothing), the rest the batch's rays. One
// persistent launch traces both, so the two share one ramp-up and one
// drain; each ray's walk and result are the probe's or the batch's own.
struct FrameThenBatch {
    Frame f;
    uint32_t width, height, tiles_x, primary;
    uint2 *frame_out;
    BatchClosest batch;
    __device__ __forceinline__ bool pixel(uint32_t r, uint32_t &x, uint32_t &y) const {
        uint32_t tile = r >> 6, lane = r & 63u;
        x = (tile % tiles_x) * 8u + (lane & 7u);
        y = (tile / tiles_x) * 8u + (lane >> 3);
        return x < width && y < height;
    }
    template <class Walk> __device__ __forceinline__ bool begin(const DeviceScene &sc, Walk &w, uint32_t r) {
        if (r >= primary) return batch.begin(sc, w, r - primary);
        uint32_t x, y;
        if (!pixel(r, x, y)) return false;
        float ox, oy, oz, dx, dy, dz;
        probe_ray(f, x, y, ox, oy, oz, dx, dy, dz);
        if (w.begin(sc, ox, oy, oz, dx, dy, dz, 3.0e38f)) return true;
        frame_out[(size_t)y * width + x] = make_uint2(kNone, __float_as_uint(w.hit.t));
        return false;
    }
    template <class Walk> __device__ __forceinline__ void finish(const DeviceScene &sc, Walk &w, uint32_t r) {
        if (r >= primary) {
            batch.finish(sc, w, r - primary);
            return;
        }
        uint32_t x, y;
        pixel(r, x, y);
        uint32_t face = w.hit.slot == kNone ? kNone : sc.faces[w.hit.slot];
        frame_out[(size_t)y * width + x] = make_uint2(face, __float_as_uint(w.hit.t));
    }
};

template <int W, int S, int G>
__global__ __launch_bounds__(G) void frame_closest_stream_kernel(DeviceScene sc, FrameThenBatch src, uint32_t n,
                                                                 uint32_t refill) {
    __shared__ uint32_t lds[S * G];
    uint32_t thread = blockIdx.x * G + threadIdx.x;
    ClosestWalk<W, S> walk;
    walk.stack = LaneStack<S>{lds + threadIdx.x, sc.spill + thread, sc.spill_stride, 0};
    stream_rays<G>(sc, walk, src, n, refill);
}

template <int W, int S, int G>
__global__ __launch_bounds__(G) void any_stream_kernel(DeviceScene sc, const float4 *rays, uint32_t n, uint32_t *out,
                                                       uint32_t refill) {
    __shared__ uint32_t lds[S * G];
    uint32_t thread = blockIdx.x * G + threadIdx.x;
    AnyWalk<W, S> walk;
    walk.stack = LaneStack<S>{lds + threadIdx.x, sc.spill + thread, sc.spill_stride, 0};
    BatchAny src{rays, out};
    stream_rays<G>(sc, walk, src, n, refill);
}

namespace {

// Workgroups of a persistent launch: as many as are resident at once, and no
// more than the items need. The residency is asked of the runtime once per
// kernel (each variant has its own registers and LDS) and device.
uint32_t resident_blocks(const void *kernel, uint32_t group, uint64_t items) {
    static std::mutex lock;
    static std::map<std::pair<int, const void *>, std::pair<int, int>> cache;  // (device, kernel) -> (per CU, CUs)
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::pair<int, int> occ;
    {
        std::lock_guard<std::mutex> g(lock);
        auto it = cache.find({dev, kernel});
        if (it == cache.end()) {
            int cus = 0, b = 0;
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kernel, (int)group, 0) != hipSuccess || b < 1) b = 1;
            it = cache.emplace(std::make_pair(dev, kernel), std::make_pair(b, cus > 0 ? cus : 1)).first;
        }
        occ = it->second;
    }
    uint64_t need = (items + group - 1) / group;
    uint64_t resident = (uint64_t)occ.first * (uint64_t)occ.second;
    return (uint32_t)(need < resident ? need : resident);
}

template <int W, int S, int G> int launch_probe_t(const DeviceScene &scene, const float *frame, uint32_t w, uint32_t h,
                                                   uint2 *out, uint32_t refill, hipStream_t stream) {
    // Primary rays keep one ray per thread whatever `refill` says: their 8x8
    // tiles are coherent, and a shared counter would scatter them over the
    // XCDs' separate L2s.
    (void)refill;
    Frame f;
    for (int k = 0; k < 16; k++) f.v[k] = frame[k];
    uint32_t blocks = (uint32_t)(launch_threads(probe_items(w, h), G) / G);
    if (blocks == 0) return 0;
    DeviceScene sc = scene;
    sc.spill_stride = blocks * G;
    hipLaunchKernelGGL((probe_kernel<W, S, G>), dim3(blocks), dim3(G), 0, stream, sc, f, w, h, out);
    return (int)hipGetLastError();
}

template <int W, int S, int G> int launch_closest_t(const DeviceScene &scene, const float4 *rays, uint32_t n,
                                                     uint2 *out, uint32_t refill, hipStream_t stream) {
    if (n == 0) return 0;
    if (refill && scene.work && n < 0x80000000u) {
        uint32_t blocks = resident_blocks((const void *)closest_stream_kernel<W, S, G>, G, n);
        DeviceScene sc = scene;
        sc.spill_stride = blocks * G;
        hipLaunchKernelGGL((closest_stream_kernel<W, S, G>), dim3(blocks), dim3(G), 0, stream, sc, rays, n, out,
                           refill);
        return (int)hipGetLastError();
    }
    uint32_t blocks = (n + G - 1) / G;
    DeviceScene sc = scene;
    sc.spill_stride = blocks * G;
    hipLaunchKernelGGL((closest_kernel<W, S, G>), dim3(blocks), dim3(G), 0, stream, sc, rays, n, out);
    return (int)hipGetLastError();
}

template <int W, int S, int G>
int launch_closest_persistent_t(const DeviceScene &scene, const float4 *rays, uint32_t n, uint2 *out, uint32_t refill,
                                uint32_t max_blocks, hipStream_t stream) {
    if (n == 0) return 0;
    uint32_t blocks = resident_blocks((const void *)closest_stream_kernel<W, S, G>, G, n);
    if (blocks > max_blocks) blocks = max_blocks;
    if (blocks == 0) return (int)hipErrorInvalidValue;
    DeviceScene sc = scene;
    sc.spill_stride = blocks * G;
    hipLaunchKernelGGL((closest_stream_kernel<W, S, G>), dim3(blocks), dim3(G), 0, stream, sc, rays, n, out, refill);
    return (int)hipGetLastError();
}

template <int W, int S, int G> int launch_any_t(const DeviceScene &scene, const float4 *rays, uint32_t n, uint32_t *out,
                                                 uint32_t refill, hipStream_t stream) {
    if (n == 0) return 0;
    if (refill && scene.work && n < 0x80000000u) {
        uint32_t blocks = resident_blocks((const void *)any_stream_kernel<W, S, G>, G, n);
        DeviceScene sc = scene;
        sc.spill_stride = blocks * G;
        hipLaunchKernelGGL((any_stream_kernel<W, S, G>), dim3(blocks), dim3(G), 0, stream, sc, rays, n, out, refill);
        return (int)hipGetLastError();
    }
    uint32_t blocks = (n + G - 1) / G;
    DeviceScene sc = scene;
    sc.spill_stride = blocks * G;
    hipLaunchKernelGGL((any_kernel<W, S, G>), dim3(blocks), dim3(G), 0, stream, sc, rays, n, out);
    return (int)hipGetLastError();
}

// The frame's primary rays and the batch in one persistent launch (refill
// above 0 and the scene's counters), or the probe then the batch.
template <int W, int S, int G>
int launch_frame_closest_t(const DeviceScene &scene, const float *frame, uint32_t w, uint32_t h, uint2 *frame_out,
                           const float4 *rays, uint32_t n, uint2 *out, uint32_t refill, hipStream_t stream) {
    uint64_t primary = probe_items(w, h), items = primary + n;
    if (!refill || !scene.work || items >= 0x80000000ull) {
        int e = launch_probe_t<W, S, G>(scene, frame, w, h, frame_out, refill, stream);
        return e ? e : launch_closest_t<W, S, G>(scene, rays, n, out, refill, stream);
    }
    FrameThenBatch src;
    for (int k = 0; k < 16; k++) src.f.v[k] = frame[k];
    src.width = w, src.height = h, src.tiles_x = (w + 7u) / 8u, src.primary = (uint32_t)primary;
    src.frame_out = frame_out;
    src.batch = BatchClosest{rays, out};
    uint32_t blocks = resident_blocks((const void *)frame_closest_stream_kernel<W, S, G>, G, items);
    DeviceScene sc = scene;
    sc.spill_stride = blocks * G;
    hipLaunchKernelGGL((frame_closest_stream_kernel<W, S, G>), dim3(blocks), dim3(G), 0, stream, sc, src,
                       (uint32_t)items, refill);
    return (int)hipGetLastError();
}

template <int W, int S, int G> uint32_t resident_t(uint64_t items, int any) {
    return any ? resident_blocks((const void *)any_stream_kernel<W, S, G>, G, items)
               : resident_blocks((const void *)closest_stream_kernel<W, S, G>, G, items);
}

template <int W, int S, int G> void fill(TraceVariant &v) {
    v.resident = &resident_t<W, S, G>;
    v.closest_persistent = &launch_closest_persistent_t<W, S, G>;
    v.probe = &launch_probe_t<W, S, G>;
    v.closest = &launch_closest_t<W, S, G>;
    v.frame_closest = &launch_frame_closest_t<W, S, G>;
    v.any = &launch_any_t<W, S, G>;
}

template <int W, int S> void fill_g(TraceVariant *table, int wi, int si) {
    fill<W, S, 64>(table[(wi * kStackSizes + si) * 3 + 0]);
    fill<W, S, 128>(table[(wi * kStackSizes + si) * 3 + 1]);
    fill<W, S, 256>(table[(wi * kStackSizes + si) * 3 + 2]);
}

template <int W> void fill_s(TraceVariant *table, int wi) {
    fill_g<W, 16>(table, wi, 0);
    fill_g<W, 24>(table, wi, 1);
    fill_g<W, 32>(table, wi, 2);
    fill_g<W, 48>(table, wi, 3);
}

}  // namespace

const TraceVariant *trace_variants() {
    static TraceVariant table[kTraceVariants];
    static bool ready = false;
    if (!ready) {
        fill_s<2>(table, 0);
        fill_s<4>(table, 1);
        fill_s<8>(table, 2);
        fill_s<kQuantizedLayout + 4>(table, 3);
        fill_s<kQuantizedLayout + 8>(table, 4);
        ready = true;
    }
    return table;
}

int variant_index(uint32_t width, uint32_t stack, uint32_t group) {
    int wi = layout_index(width);
    int si = stack == 16 ? 0 : stack == 24 ? 1 : stack == 32 ? 2 : stack == 48 ? 3 : -1;
    int gi = group == 64 ? 0 : group == 128 ? 1 : group == 256 ? 2 : -1;
    if (wi < 0 || si < 0 || gi < 0) return -1;
    return (wi * kStackSizes + si) * 3 + gi;
}

}  // namespace lr
