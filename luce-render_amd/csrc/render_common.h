// Device-side common ground of the wavefront integrator: launch parameters,
// the constants rows, the path state, queues, the sampler and colour science.
//
// Layouts the host packs (luce_render/loading.py) and the kernels read are the
// reference's, so a scene packs identically for either:
//   constants  32 float4 rows, K_* below (shaders/common.glsl:79-86 and the
//              rows named in lights.glsl, lighttree.glsl, guide.glsl,
//              openpbr.glsl; filled as src/render/loading.lucb:253-284 does);
//   lights     7 float4 per light, then the light tree's 3 float4 per node
//              (src/render/lights.lucb, light_tree.lucb);
//   materials  16 float4 per material (src/render/materials.lucb:105-119);
//   emitters   2 u32 per tree node, 4 per emitter, then each triangle's
//              emitter (src/render/light_tree.lucb:16-20).
// Path state is field-major (field * pool + path) as in common.glsl:89-162:
// every lane of a wave touches consecutive addresses of one field.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "guide_net.h"
#include "trace_device.h"

namespace lr {
namespace rt {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInf = 3.0e38f;
constexpr uint32_t kNoneU = 0xffffffffu;

// Constants rows.
constexpr int K_INFO = 0;         // x lights, y max bounces, z indirect clamp, w cut-outs present
constexpr int K_BACKGROUND_FIT = 1;
constexpr int K_BACKGROUND_RGB = 2;
constexpr int K_FILM = 3;         // 3 rows: XYZ under E to ACEScg
constexpr int K_CMF = 6;          // the colour-matching integrals
constexpr int K_LIMITS = 7;       // diffuse, glossy, transmission bounce limits
constexpr int K_TREE = 8;         // x nodes, y first float4 of the nodes in lights[], z distant chance, w emitters
constexpr int K_DISPLAY = 9;      // x exposure (stops), y view (0 standard, 1 neutral)
constexpr int K_EMITTERS = 10;    // x emitter records' first u32, y triangle map's first u32
constexpr int K_ADAPTIVE = 11;
constexpr int K_GUIDE = 12;       // x on, y cell width per unit distance, z guide share, w samples before guiding
constexpr int K_GUIDE_SIZE = 13;  // x smallest cell width
constexpr int K_CAMERA = 14;      // x light paths a sample, y 1/pixel area at unit distance, z lens radius, w focus
constexpr int K_VIEW = 15;        // camera forward axis
constexpr int K_DOME = 16;        // x the light whose image this is (-1: none), y its texels' first float4, z width, w height
constexpr int K_DOME_CDF = 17;    // x its cumulative weights' first float, y the weights' sum
constexpr int K_BASIS = 18;       // 3 rows
constexpr int K_EMIT = 21;        // x chance a light path starts afar, y disk radius
constexpr int K_EMIT_CENTER = 22;

// State fields (float4 and u32), as common.glsl:89-157.
constexpr uint32_t SV_ORIGIN = 0, SV_DIRECTION = 1, SV_THROUGHPUT = 2, SV_RADIANCE = 3, SV_LAMBDA = 4,
                   SV_SHADOW_ORIGIN = 5, SV_SHADOW_DIRECTION = 6, SV_SHADOW_RADIANCE = 7, SV_HIT = 8, SV_VERTEX = 9,
                   SV_GUIDE0 = 10, SV_MIS = 13,
                   SV_GUIDE_POS0 = 14;  // learned guide: (point, octahedral geometric normal bits) per guide record
constexpr uint32_t SU_PIXEL = 0, SU_BOUNCE = 1, SU_HIT = 2, SU_NORMAL = 3, SU_GUIDE_CELL0 = 4, SU_GUIDE_DIR0 = 7,
                   SU_SPLAT = 10;
constexpr uint32_t kStateV = 17, kStateU = 11;

// Queues: counters [0..16), then each queue's paths from QUEUE_BASE.
constexpr uint32_t Q_CLOSEST = 0, Q_SURFACE = 2, Q_MISS = 3, Q_SHADOW = 4, Q_LIGHT = 5;
constexpr uint32_t kQueues = 6, kQueueBase = 48;

constexpr uint32_t kLightHit = 0x80000000u;
constexpr uint32_t kAfarEmitter = 0x40000000u;

// Film planes (common.glsl:64-75).
constexpr uint32_t FILM_SUM = 0, FILM_HALF = 1, FILM_SCRATCH = 2, FILM_SPLAT = 3;

// What changes from pass to pass, kept on the device so a recorded pass
// replays unchanged (begin_pass / end_pass advance it).
struct PassState {
    uint32_t sample;        // this pass's sample index
    uint32_t band;          // this pass's band
    uint32_t band_start;
    uint32_t band_count;
    uint32_t light_count;   // light paths in this pass
    uint32_t light_samples; // light-tracing samples taken so far (film_convert divides splats by them)
    uint32_t samples_done;  // whole samples finished
    uint32_t bands;         // bands a sample
    unsigned long long closest_rays;  // rays the closest-hit kernels have traced
    unsigned long long shadow_rays;   // shadow rays traced
    uint32_t sub_samples;   // samples this pass carries (one band a sample: 1 to Params::pass_samples)
    uint32_t unused;
};

// Launch parameters fixed for a view: pointers, sizes, the frame.
// A node of the analytic lights' culling tree: a box holding its lights
// (grown outward from their exact bounds) and `light`, the light of a leaf
// (kNoneU for an interior node or an empty leaf, whose box holds nothing).
struct LightNode {
    float lo[3];
    uint32_t unused;
    float hi[3];
    uint32_t light;
};

struct Params {
    float frame[16];
    uint32_t pool;        // path slots: the stride of every state field
    uint32_t seed;
    uint32_t pixels;
    uint32_t width, height;
    uint32_t band_size;   // camera paths a band
    uint32_t light_paths; // light paths a sample (in the first band)
    uint32_t spectral;
    // Samples a pass may carry when a sample is one band (its slots hold that
    // many samples' camera paths, then their light paths), and the render's
    // sample target, which the last pass's samples stop at.
    uint32_t pass_samples;
    uint32_t sample_target;
    const float4 *lights;
    const float4 *constants;
    const float4 *materials;
    const uint32_t *emitters;
    const float *tables;
    const uint4 *shading;
    const float *uvs;
    // Materials' node-graph programs and the textures their image nodes read
    // (texture_table rows: first texel, width, height, 0; textures: RGBA
    // texels, rows from the top). Empty when no material has a program.
    const uint32_t *programs;
    const float4 *textures;
    const uint4 *texture_table;
    uint32_t program_words;
    uint32_t texture_count;
    // Gaussian splats (the scene's): records (4 uint4 each, leaf order),
    // harmonics words, and their binary tree (two float4 a node: lo + first,
    // hi + count). splat_count 0: no splats.
    const uint4 *splat_records;
    const uint32_t *splat_words;
    const float4 *splat_nodes;
    uint32_t splat_count;
    float4 *state_v;
    uint32_t *state_u;
    uint32_t *queues;
    float4 *film;
    float4 *image;
    float4 *display;
    uint8_t *guide;        // guiding grid (keys, lobes, stats); null without guiding
    // The analytic lights' culling tree: a complete binary tree in
    // breadth-first order from index 1 (node k's children at 2k and 2k + 1,
    // so two siblings are one 64-byte read; leaves from light_nodes / 2), and
    // light_nodes, twice its leaves (0 when no light can be hit).
    const LightNode *light_tree;
    uint32_t light_nodes;
    // The distant and dome lights (nothing to hit, found by direction), in
    // index order: what the miss shading and the distant pick walk.
    const uint32_t *far_lights;
    uint32_t far_count;
    // The learned guide (guiding 2): the network, the per-slot mixtures the
    // inference writes each bounce (two rows of `pool` 16-byte words: lobes
    // 0-1, then 2-3, each an octahedral mean and a (kappa, weight) pair of
    // halves), the training gradient rows and schedule.
    net::NetBuffers net;
    uint4 *guide_mix;
    float *net_partials;
    uint32_t learned;
    uint32_t train_samples;  // samples whose paths train it
    uint32_t train_batch;    // records a training step draws
    float train_rate;
    PassState *pass;
    DeviceScene scene;
};

// ------------------------------------------------------ a pass's samples
// A pass carries sub_samples samples of its band: camera slots j * band + i
// (sample `sample + j`, the band's i-th camera path), then light slots
// sub_samples * band + j * lights + i (sample `sample + j`, the i-th light
// path). Every path draws its numbers by its sample and by the slot it would
// hold in a pass of one sample, so a pass of several samples renders what
// that many passes of one render.

__device__ __forceinline__ uint32_t pass_cameras(const PassState &ps) { return ps.band_count * ps.sub_samples; }

struct SlotIdentity {
    uint32_t sample;  // the path's sample index
    uint32_t slot;    // its slot in a pass of one sample
};

__device__ __forceinline__ SlotIdentity slot_identity(const PassState &ps, uint32_t path) {
    if (ps.sub_samples <= 1u) return {ps.sample, path};
    uint32_t cameras = pass_cameras(ps);
    if (path < cameras) {
        uint32_t j = path / ps.band_count;
        return {ps.sample + j, path - j * ps.band_count};
    }
    uint32_t i = path - cameras;
    uint32_t j = ps.light_count ? i / ps.light_count : 0u;
    return {ps.sample + j, ps.band_count + (i - j * ps.light_count)};
}

// ------------------------------------------------------------- state access

struct State {
    const Params &p;
    __device__ State(const Params &params) : p(params) {}
    __device__ __forceinline__ float4 v(uint32_t field, uint32_t path) const {
        return p.state_v[(size_t)field * p.pool + path];
    }
    __device__ __forceinline__ void set_v(uint32_t field, uint32_t path, float4 x) const {
        p.state_v[(size_t)field * p.pool + path] = x;
    }
    __device__ __forceinline__ uint32_t u(uint32_t field, uint32_t path) const {
        return p.state_u[(size_t)field * p.pool + path];
    }
    __device__ __forceinline__ void set_u(uint32_t field, uint32_t path, uint32_t x) const {
        p.state_u[(size_t)field * p.pool + path] = x;
    }
};

__device__ __forceinline__ float4 K(const Params &p, int row) { return p.constants[row]; }

// --------------------------------------------------------------- wave queues

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Staged ids per queue: up to eight iterations of a workgroup between flushes.
constexpr uint32_t kStageIterations = 8u;

// Queue pushes of one workgroup staged in LDS: each wave claims LDS slots with
// one LDS atomic, and a flush claims the workgroup's run of every queue with
// one global atomic per queue and copies the staged ids out in order. Global
// atomics on one counter from every XCD serialize at the memory side, so one
// per wave per push was the integrator's largest cost; staging several
// iterations of a workgroup cuts them by that factor. Q queues of CAP ids.
template <uint32_t Q, uint32_t CAP> struct StagedPushes {
    uint32_t ids[Q][CAP];
    uint32_t count[Q];
    uint32_t base[Q];
    uint32_t queue[Q];

    // Every thread of the workgroup calls it before the loop.
    __device__ __forceinline__ void init(const uint32_t (&queues)[Q]) {
        if (threadIdx.x < Q) {
            count[threadIdx.x] = 0u;
            queue[threadIdx.x] = queues[threadIdx.x];
        }
        __syncthreads();
    }

    // Wave-wide: the active lanes' paths go to staged queue q.
    __device__ __forceinline__ void push(uint32_t q, uint32_t path, bool active) {
        uint64_t mask = __ballot(active);
        if (mask == 0) return;
        uint32_t lane = lane_id();
        uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
        uint32_t at = 0;
        if (lane == leader) at = atomicAdd(&count[q], (uint32_t)__popcll(mask));
        at = __shfl(at, (int)leader);
        uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
        if (active) ids[q][at + rank] = path;
    }

    // Workgroup-wide, once per loop iteration (block-uniform): writes the
    // staged ids out when `last` or when another iteration might not fit.
    __device__ __forceinline__ void sync(const Params &p, bool last) {
        __syncthreads();
        bool flush = last;
        for (uint32_t q = 0; q < Q; q++) flush = flush || count[q] + blockDim.x > CAP;
        if (!flush) return;
        if (threadIdx.x < Q) {
            uint32_t c = count[threadIdx.x];
            base[threadIdx.x] = c ? atomicAdd(&p.queues[queue[threadIdx.x]], c) : 0u;
        }
        __syncthreads();
        for (uint32_t q = 0; q < Q; q++) {
            uint32_t c = count[q];
            uint32_t *dst = p.queues + kQueueBase + (size_t)queue[q] * p.pool + base[q];
            for (uint32_t i = threadIdx.x; i < c; i += blockDim.x) dst[i] = ids[q][i];
        }
        __syncthreads();
        if (threadIdx.x < Q) count[threadIdx.x] = 0u;
        __syncthreads();
    }
};

// Per-bounce grids are sized for every slot; once few paths remain most of
// their workgroups have no item. With one request such a workgroup leaves
// after one load of the queue's count, before the row plan (workgroup 0
// stays: kernels clear counters there). It would have run no iteration.
template <class F> __device__ __forceinline__ bool idle_block(const Params *plist, uint32_t n, F count_of) {
    if (n != 1u || blockIdx.x == 0u) return false;
    return blockIdx.x * blockDim.x >= count_of(plist[0]);
}

// A launch over several renders (rows of a device array of Params): every
// workgroup first lays the rows' item counts end to end, each padded to whole
// workgroups, so that one grid of workgroups sized for the machine covers all
// rows and every workgroup iteration works for exactly one row. A shared
// launch then costs no more workgroups than the work needs, however uneven
// the rows are. At most kMaxRows rows.
constexpr uint32_t kMaxRows = 128u;

struct RowPlan {
    uint32_t off[kMaxRows + 1];  // padded prefix: row r's items start at off[r]
    uint32_t count[kMaxRows];

    // Workgroup-wide: count_of(p) is a row's item count. Returns the padded total.
    template <class F> __device__ __forceinline__ uint32_t make(const Params *plist, uint32_t n, F count_of) {
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            uint32_t c = count_of(plist[i]);
            count[i] = c;
            off[i + 1] = (c + blockDim.x - 1u) / blockDim.x * blockDim.x;
        }
        __syncthreads();
        if (threadIdx.x < 64u) {
            // one wave scans the padded counts, 64 at a time
            uint32_t lane = threadIdx.x, carry = 0;
            for (uint32_t base = 0; base < n; base += 64u) {
                uint32_t i = base + lane;
                uint32_t v = i < n ? off[i + 1] : 0u;
                for (uint32_t d = 1; d < 64u; d <<= 1) {
                    uint32_t t = __shfl_up(v, d);
                    if (lane >= d) v += t;
                }
                if (i < n) off[i + 1] = v + carry;
                carry += __shfl(v, 63);
            }
            if (lane == 0) off[0] = 0u;
        }
        __syncthreads();
        return off[n];
    }
    // The row whose padded range holds item g (workgroup-uniform).
    __device__ __forceinline__ uint32_t row_of(uint32_t g, uint32_t n) const {
        uint32_t lo = 0, hi = n;  // off[lo] <= g < off[hi]
        while (hi - lo > 1u) {
            uint32_t mid = (lo + hi) >> 1;
            if (off[mid] <= g) lo = mid;
            else hi = mid;
        }
        return __builtin_amdgcn_readfirstlane(lo);
    }
};

// The items of every row: g walks the padded ranges a workgroup at a time
// (workgroup-uniform), r is g's row, `at` the item within the row, `active`
// whether it exists. `last` tells whether the next iteration leaves row r.
#define FOR_ROW_ITEMS(plan, n, total, g, r, at, active)                                                    \
    for (uint32_t g = blockIdx.x * blockDim.x; g < (total); g += gridDim.x * blockDim.x)                  \
        for (uint32_t r = (plan).row_of(g, n), _once = 1; _once; _once = 0)                                \
            for (uint32_t at = g - (plan).off[r] + threadIdx.x; _once; _once = 0)                          \
                for (bool active = at < (plan).count[r]; _once; _once = 0)

__device__ __forceinline__ bool row_ends(const RowPlan &plan, uint32_t n, uint32_t total, uint32_t g, uint32_t r) {
    uint32_t next = g + gridDim.x * blockDim.x;
    return next >= total || plan.row_of(next, n) != r;
}

__device__ __forceinline__ uint32_t queue_path(const Params &p, uint32_t queue, uint32_t at) {
    return p.queues[kQueueBase + (size_t)queue * p.pool + at];
}

// ------------------------------------------------------------------- vectors

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 v3(float4 a) { return V3{a.x, a.y, a.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) {
    float s = a.x * b.x;
    s = s + a.y * b.y;
    return s + a.z * b.z;
}
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 normalize(V3 a) { return a / length(a); }
__device__ __forceinline__ float distance(V3 a, V3 b) { return length(a - b); }
__device__ __forceinline__ float4 f4(V3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
__device__ __forceinline__ V3 reflect(V3 i, V3 n) { return i - n * (2.0f * dot(n, i)); }
// GLSL refract: zero when totally internally reflected.
__device__ __forceinline__ V3 refract(V3 i, V3 n, float eta) {
    float d = dot(n, i);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return V3{0.0f, 0.0f, 0.0f};
    return i * eta - n * (eta * d + sqrtf(k));
}

// Spectral values: four wavelengths, or ACEScg in xyz.
using Spec = float4;
__device__ __forceinline__ Spec spec(float s) { return make_float4(s, s, s, s); }
__device__ __forceinline__ Spec operator+(Spec a, Spec b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
__device__ __forceinline__ Spec operator-(Spec a, Spec b) { return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
__device__ __forceinline__ Spec operator*(Spec a, Spec b) { return make_float4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }
__device__ __forceinline__ Spec operator*(Spec a, float s) { return make_float4(a.x * s, a.y * s, a.z * s, a.w * s); }
__device__ __forceinline__ Spec operator*(float s, Spec a) { return a * s; }
__device__ __forceinline__ Spec operator/(Spec a, Spec b) { return make_float4(a.x / b.x, a.y / b.y, a.z / b.z, a.w / b.w); }
__device__ __forceinline__ Spec operator/(Spec a, float s) { return make_float4(a.x / s, a.y / s, a.z / s, a.w / s); }
__device__ __forceinline__ Spec smin(Spec a, Spec b) {
    return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), fminf(a.w, b.w));
}
__device__ __forceinline__ Spec sclamp(Spec a, float lo, float hi) {
    return make_float4(fminf(fmaxf(a.x, lo), hi), fminf(fmaxf(a.y, lo), hi), fminf(fmaxf(a.z, lo), hi),
                       fminf(fmaxf(a.w, lo), hi));
}
__device__ __forceinline__ float lane(Spec s, int i) { return i == 0 ? s.x : i == 1 ? s.y : i == 2 ? s.z : s.w; }
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ Spec mixs(Spec a, Spec b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float clampf(float x, float a, float b) { return fminf(fmaxf(x, a), b); }
__device__ __forceinline__ float fractf(float x) { return x - floorf(x); }

// ------------------------------------------------------------------- sampler
// Owen-scrambled Sobol by Burley's hashing (shaders/sampler.glsl:8-55):
// integer arithmetic, so every build and the oracle draw identical numbers.

__constant__ static const uint32_t kSobol[128] = {
    0x80000000u, 0x40000000u, 0x20000000u, 0x10000000u, 0x08000000u, 0x04000000u, 0x02000000u, 0x01000000u,
    0x00800000u, 0x00400000u, 0x00200000u, 0x00100000u, 0x00080000u, 0x00040000u, 0x00020000u, 0x00010000u,
    0x00008000u, 0x00004000u, 0x00002000u, 0x00001000u, 0x00000800u, 0x00000400u, 0x00000200u, 0x00000100u,
    0x00000080u, 0x00000040u, 0x00000020u, 0x00000010u, 0x00000008u, 0x00000004u, 0x00000002u, 0x00000001u,
    0x80000000u, 0xc0000000u, 0xa0000000u, 0xf0000000u, 0x88000000u, 0xcc000000u, 0xaa000000u, 0xff000000u,
    0x80800000u, 0xc0c00000u, 0xa0a00000u, 0xf0f00000u, 0x88880000u, 0xcccc0000u, 0xaaaa0000u, 0xffff0000u,
    0x80008000u, 0xc000c000u, 0xa000a000u, 0xf000f000u, 0x88008800u, 0xcc00cc00u, 0xaa00aa00u, 0xff00ff00u,
    0x80808080u, 0xc0c0c0c0u, 0xa0a0a0a0u, 0xf0f0f0f0u, 0x88888888u, 0xccccccccu, 0xaaaaaaaau, 0xffffffffu,
    0x80000000u, 0xc0000000u, 0x60000000u, 0x90000000u, 0xe8000000u, 0x5c000000u, 0x8e000000u, 0xc5000000u,
    0x68800000u, 0x9cc00000u, 0xee600000u, 0x55900000u, 0x80680000u, 0xc09c0000u, 0x60ee0000u, 0x90550000u,
    0xe8808000u, 0x5cc0c000u, 0x8e606000u, 0xc5909000u, 0x6868e800u, 0x9c9c5c00u, 0xeeee8e00u, 0x5555c500u,
    0x8000e880u, 0xc0005cc0u, 0x60008e60u, 0x9000c590u, 0xe8006868u, 0x5c009c9cu, 0x8e00eeeeu, 0xc5005555u,
    0x80000000u, 0xc0000000u, 0x20000000u, 0x50000000u, 0xf8000000u, 0x74000000u, 0xa2000000u, 0x93000000u,
    0xd8800000u, 0x25400000u, 0x59e00000u, 0xe6d00000u, 0x78080000u, 0xb40c0000u, 0x82020000u, 0xc3050000u,
    0x208f8000u, 0x51474000u, 0xfbea2000u, 0x75d93000u, 0xa0858800u, 0x914e5400u, 0xdbe79e00u, 0x25db6d00u,
    0x58800080u, 0xe54000c0u, 0x79e00020u, 0xb6d00050u, 0x800800f8u, 0xc00c0074u, 0x200200a2u, 0x50050093u};

__device__ __forceinline__ uint32_t hash_u32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x21f0aaadu;
    x ^= x >> 15;
    x *= 0x735a2d97u;
    x ^= x >> 15;
    return x;
}
__device__ __forceinline__ uint32_t hash_combine(uint32_t seed, uint32_t value) {
    return seed ^ (hash_u32(value) + 0x9e3779b9u + (seed << 6) + (seed >> 2));
}
__device__ __forceinline__ uint32_t owen(uint32_t x, uint32_t seed) {
    x = __builtin_bitreverse32(x);
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return __builtin_bitreverse32(x);
}
// The first four Sobol dimensions of `index` (sampler.glsl:29-35: the XOR of
// kSobol's columns for the index's set bits), without a per-lane loop over
// the bits or table loads:
//  - dimension 0's matrix is the bit reversal;
//  - dimension 1's is Pascal's triangle mod 2, so by Lucas' theorem output
//    bit 31-j is the XOR of the index bits b whose bit pattern holds j's
//    (j & b == j): a superset sum over the five bits of a bit position, five
//    shift-mask-XOR steps, then the reversal;
//  - dimensions 2 and 3 take one pass over the 32 bit positions, each set bit
//    a mask (no branch), unrolled by 8 with the columns read at wave-uniform
//    addresses (fully unrolled, its immediates doubled the shading kernels'
//    spills: 2.38 against 2.12 ms a 1080p sample of the glossy workload).
// Equal to the table loop for every index (any XOR order is the same word).
__device__ __forceinline__ void sobol4(uint32_t index, uint32_t (&x)[4]) {
    x[0] = __builtin_bitreverse32(index);
    uint32_t g = index;
    g ^= (g >> 1) & 0x55555555u;
    g ^= (g >> 2) & 0x33333333u;
    g ^= (g >> 4) & 0x0F0F0F0Fu;
    g ^= (g >> 8) & 0x00FF00FFu;
    g ^= (g >> 16) & 0x0000FFFFu;
    x[1] = __builtin_bitreverse32(g);
    uint32_t x2 = 0u, x3 = 0u, rest = index;
#pragma unroll 8
    for (uint32_t bit = 0; bit < 32u; bit++, rest >>= 1) {
        uint32_t m = 0u - (rest & 1u);
        x2 ^= m & kSobol[64u + bit];
        x3 ^= m & kSobol[96u + bit];
    }
    x[2] = x2;
    x[3] = x3;
}

constexpr uint32_t GROUP_CAMERA = 0, GROUP_WAVELENGTH = 1;
__device__ __forceinline__ uint32_t group_light(uint32_t b) { return 2u + 3u * b; }
__device__ __forceinline__ uint32_t group_bsdf(uint32_t b) { return 3u + 3u * b; }
__device__ __forceinline__ uint32_t group_opacity(uint32_t b) { return 4u + 3u * b; }

__device__ __forceinline__ float4 sample4(uint32_t seed0, uint32_t pixel, uint32_t index, uint32_t group) {
    uint32_t seed = hash_combine(hash_combine(seed0, pixel), group);
    uint32_t shuffled = owen(index, hash_u32(seed));
    uint32_t x[4];
    sobol4(shuffled, x);
    float r[4];
#pragma unroll
    for (uint32_t d = 0; d < 4; d++) r[d] = (float)(owen(x[d], hash_combine(seed, d + 1u)) >> 8) * (1.0f / 16777216.0f);
    return make_float4(r[0], r[1], r[2], r[3]);
}

// ------------------------------------------------------------------ helpers

// Shirley–Chiu concentric disk (common.glsl:103-110).
__device__ __forceinline__ float2 concentric(float ux, float uy) {
    float ox = 2.0f * ux - 1.0f, oy = 2.0f * uy - 1.0f;
    if (ox == 0.0f && oy == 0.0f) return make_float2(0.0f, 0.0f);
    float r, theta;
    if (fabsf(ox) > fabsf(oy)) {
        r = ox;
        theta = (kPi / 4.0f) * (oy / ox);
    } else {
        r = oy;
        theta = (kPi / 2.0f) - (kPi / 4.0f) * (ox / oy);
    }
    return make_float2(r * cosf(theta), r * sinf(theta));
}

// Duff et al.'s orthonormal basis (common.glsl:270-276).
__device__ __forceinline__ void basis(V3 n, V3 &t, V3 &b) {
    float s = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (s + n.z);
    float c = n.x * n.y * a;
    t = V3{1.0f + s * n.x * n.x * a, s * c, -s * n.x};
    b = V3{c, s + n.y * n.y * a, -n.y};
}

__device__ __forceinline__ V3 offset_origin(V3 p, V3 n) {
    float scale = fmaxf(1.0f, fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
    return p + n * (1e-4f * scale);
}

// snorm16 x 2 octahedral (common.glsl:256-267; GLSL's pack/unpackSnorm2x16).
__device__ __forceinline__ uint32_t octahedral_encode(V3 n) {
    float l = fabsf(n.x) + fabsf(n.y) + fabsf(n.z);
    float px = n.x / l, py = n.y / l;
    if (n.z < 0.0f) {
        float qx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
        float qy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
        px = qx;
        py = qy;
    }
    int ix = (int)rintf(clampf(px, -1.0f, 1.0f) * 32767.0f);
    int iy = (int)rintf(clampf(py, -1.0f, 1.0f) * 32767.0f);
    return ((uint32_t)ix & 0xffffu) | (((uint32_t)iy & 0xffffu) << 16);
}
__device__ __forceinline__ V3 octahedral_decode(uint32_t packed) {
    float x = fmaxf(-1.0f, (float)(int16_t)(packed & 0xffffu) / 32767.0f);
    float y = fmaxf(-1.0f, (float)(int16_t)(packed >> 16) / 32767.0f);
    V3 n{x, y, 1.0f - fabsf(x) - fabsf(y)};
    if (n.z < 0.0f) {
        float nx = (1.0f - fabsf(n.y)) * (n.x >= 0.0f ? 1.0f : -1.0f);
        float ny = (1.0f - fabsf(n.x)) * (n.y >= 0.0f ? 1.0f : -1.0f);
        n.x = nx;
        n.y = ny;
    }
    return normalize(n);
}

// Power-heuristic weight from the other two techniques' density ratios
// (common.glsl:238-242).
__device__ __forceinline__ float mis_weight(float a, float b) {
    a = fminf(a, 1e18f);
    b = fminf(b, 1e18f);
    return 1.0f / (1.0f + a * a + b * b);
}

__device__ __forceinline__ uint32_t bounce_total(uint32_t b) { return b & 0xffu; }

// The bounces after one more of `lobe` (0 diffuse, 1 glossy, 2 transmission),
// kNoneU past that kind's limit (common.glsl:229-234).
__device__ __forceinline__ uint32_t next_bounce(const Params &p, uint32_t bounces, uint32_t lobe) {
    uint32_t shift = 8u * (lobe + 1u);
    uint32_t kind = ((bounces >> shift) & 0xffu) + 1u;
    float4 lim = K(p, K_LIMITS);
    float limit = lobe == 0 ? lim.x : lobe == 1 ? lim.y : lim.z;
    if ((float)kind > limit || bounce_total(bounces) >= 255u) return kNoneU;
    return (bounces & ~(0xffu << shift)) + (kind << shift) + 1u;
}

// ------------------------------------------------------------ colour science
// shaders/spectrum.glsl.

constexpr float kLambdaMin = 360.0f, kLambdaMax = 830.0f;

__device__ __forceinline__ Spec sample_wavelengths(float u) {
    float l[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        float v = fractf(u + (float)i * 0.25f);
        l[i] = 538.0f - 138.888889f * atanhf(0.85691062f - 1.82750197f * v);
    }
    return make_float4(l[0], l[1], l[2], l[3]);
}

__device__ __forceinline__ float wavelength_pdf(float lambda) {
    if (lambda < kLambdaMin || lambda > kLambdaMax) return 0.0f;
    float c = coshf(0.0072f * (lambda - 538.0f));
    return 0.0039398042f / (c * c);
}

__device__ __forceinline__ float cmf_lobe(float x, float mean, float below, float above) {
    float t = (x - mean) / (x < mean ? below : above);
    return expf(-0.5f * t * t);
}
// Wyman, Sloan and Shirley's multi-lobe CIE 1931 fit.
__device__ __forceinline__ V3 cmf(float l) {
    return V3{1.056f * cmf_lobe(l, 599.8f, 37.9f, 31.0f) + 0.362f * cmf_lobe(l, 442.0f, 16.0f, 26.7f) -
                  0.065f * cmf_lobe(l, 501.1f, 20.4f, 26.2f),
              0.821f * cmf_lobe(l, 568.8f, 46.9f, 40.5f) + 0.286f * cmf_lobe(l, 530.9f, 16.3f, 31.1f),
              1.217f * cmf_lobe(l, 437.0f, 11.8f, 36.0f) + 0.681f * cmf_lobe(l, 459.0f, 26.0f, 13.8f)};
}

// A colour as a Spec: its sigmoid fit at the wavelengths, or its RGB.
template <bool SPECTRAL> __device__ __forceinline__ Spec spec_of(float4 fit, float4 rgb, Spec lambda) {
    if constexpr (SPECTRAL) {
        float r[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float x = (lane(lambda, i) - kLambdaMin) / (kLambdaMax - kLambdaMin);
            float z = (fit.x * x + fit.y) * x + fit.z;
            r[i] = fit.w * (0.5f + z / (2.0f * sqrtf(1.0f + z * z)));
        }
        return make_float4(r[0], r[1], r[2], r[3]);
    } else {
        return make_float4(rgb.x, rgb.y, rgb.z, 0.0f);
    }
}

__device__ __forceinline__ float blackbody(float lambda, float kelvin) {
    float peak = 2.8977721e6f / kelvin;
    float x = 1.4387769e7f / (lambda * kelvin);
    float r = peak / lambda;
    return r * r * r * r * r * (142.32492f / (expf(x) - 1.0f));
}

__device__ __forceinline__ float logistic(float x) { return 1.0f / (1.0f + expf(-x)); }

template <bool SPECTRAL> __device__ __forceinline__ Spec spec_of_texel(const Params &p, V3 rgb, Spec lambda) {
    if constexpr (SPECTRAL) {
        V3 w{dot(v3(K(p, K_BASIS)), rgb), dot(v3(K(p, K_BASIS + 1)), rgb), dot(v3(K(p, K_BASIS + 2)), rgb)};
        float r[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float blue = 1.0f - logistic((lane(lambda, i) - 490.0f) / 12.0f);
            float red = logistic((lane(lambda, i) - 595.0f) / 12.0f);
            r[i] = w.x * blue + w.y * (1.0f - blue - red) + w.z * red;
        }
        return make_float4(r[0], r[1], r[2], r[3]);
    } else {
        return make_float4(rgb.x, rgb.y, rgb.z, 0.0f);
    }
}

template <bool SPECTRAL> __device__ __forceinline__ float spec_luminance(const Params &p, Spec s, Spec lambda) {
    if constexpr (SPECTRAL) {
        float sum = 0.0f;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float pdf = wavelength_pdf(lane(lambda, i));
            if (pdf > 0.0f) sum += cmf(lane(lambda, i)).y * (lane(s, i) / pdf);
        }
        return sum * 0.25f / K(p, K_CMF).y;
    } else {
        return s.x * 0.2722287168f + s.y * 0.6740817658f + s.z * 0.0536895174f;
    }
}

template <bool SPECTRAL> __device__ __forceinline__ float spec_max(Spec s) {
    if constexpr (SPECTRAL) return fmaxf(fmaxf(s.x, s.y), fmaxf(s.z, s.w));
    else return fmaxf(s.x, fmaxf(s.y, s.z));
}

template <bool SPECTRAL> __device__ __forceinline__ V3 spec_to_film(const Params &p, Spec radiance, Spec lambda) {
    if constexpr (SPECTRAL) {
        V3 sum{0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float pdf = wavelength_pdf(lane(lambda, i));
            if (pdf > 0.0f) sum = sum + cmf(lane(lambda, i)) * (lane(radiance, i) / pdf);
        }
        float4 c = K(p, K_CMF);
        return V3{sum.x * 0.25f / c.x, sum.y * 0.25f / c.y, sum.z * 0.25f / c.z};
    } else {
        return V3{radiance.x, radiance.y, radiance.z};
    }
}

}  // namespace rt
}  // namespace lr
