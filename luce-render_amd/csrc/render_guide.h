// Path guiding by a hash grid of von Mises–Fisher mixtures (shaders/guide.glsl,
// shaders/guide_fit.comp): the plain guide, which the learned guide stands in
// for and which judges it. One buffer in three regions: u32 keys, float4 lobes
// (4 lobes then the weights per cell), float stats (5 per lobe).
#pragma once

#include "render_common.h"

namespace lr {
namespace rt {

constexpr uint32_t GUIDE_CELLS = 524288u;
constexpr uint32_t GUIDE_LOBES = 4u;
constexpr uint32_t GUIDE_RING = 3u;
constexpr uint32_t GUIDE_TRAINING = 1u;
constexpr uint32_t GUIDE_PROBES = 8u;
constexpr size_t kGuideBytes = (size_t)GUIDE_CELLS * (4 + 16 * (GUIDE_LOBES + 1) + 4 * 5 * GUIDE_LOBES);

__device__ __forceinline__ bool guiding(const Params &p) {
    return (p.guide != nullptr || p.learned) && K(p, K_GUIDE).x > 0.5f;
}
// The learned guide has taken a training step (its mixtures are in use).
__device__ __forceinline__ bool net_ready(const Params &p) { return p.learned && p.net.stats[1] >= 1.0f; }
__device__ __forceinline__ uint32_t *guide_keys(const Params &p) { return reinterpret_cast<uint32_t *>(p.guide); }
__device__ __forceinline__ float4 *guide_lobes(const Params &p) {
    return reinterpret_cast<float4 *>(p.guide + (size_t)GUIDE_CELLS * 4);
}
__device__ __forceinline__ float *guide_stats(const Params &p) {
    return reinterpret_cast<float *>(p.guide + (size_t)GUIDE_CELLS * (4 + 16 * (GUIDE_LOBES + 1)));
}

// guide.glsl:101-122: the cell of point x with normal n, made on first use.
__device__ __forceinline__ uint32_t guide_cell(const Params &p, V3 x, V3 n) {
    float minimum = K(p, K_GUIDE_SIZE).x;
    V3 cam{p.frame[0], p.frame[1], p.frame[2]};
    float wanted = fmaxf(minimum, distance(x, cam) * K(p, K_GUIDE).y);
    uint32_t level = (uint32_t)clampf(ceilf(log2f(wanted / minimum)), 0.0f, 31.0f);
    float size = minimum * exp2f((float)level);
    int qx = (int)floorf(x.x / size), qy = (int)floorf(x.y / size), qz = (int)floorf(x.z / size);
    uint32_t octant = (n.x >= 0.0f ? 1u : 0u) | (n.y >= 0.0f ? 2u : 0u) | (n.z >= 0.0f ? 4u : 0u);
    uint32_t key = hash_u32((uint32_t)qx * 73856093u ^
                            hash_u32((uint32_t)qy * 19349663u ^ hash_u32((uint32_t)qz * 83492791u ^ (octant | (level << 3u)))));
    key = key > 1u ? key : 1u;
    uint32_t slot = key % GUIDE_CELLS;
    uint32_t *keys = guide_keys(p);
    for (uint32_t probe = 0; probe < GUIDE_PROBES; probe++) {
        uint32_t at = (slot + probe) % GUIDE_CELLS;
        uint32_t seen = __atomic_load_n(&keys[at], __ATOMIC_RELAXED);
        if (seen == key) return at;
        if (seen != 0u) continue;
        uint32_t found = atomicCAS(&keys[at], 0u, key);
        if (found == 0u || found == key) return at;
    }
    return kNoneU;
}

__device__ __forceinline__ float guide_count(const Params &p, uint32_t cell) {
    const float *s = guide_stats(p);
    float c = 0.0f;
    for (uint32_t k = 0; k < GUIDE_LOBES; k++) c += s[(cell * GUIDE_LOBES + k) * 5u + 4u];
    return c;
}

__device__ __forceinline__ float4 guide_lobe(const Params &p, uint32_t cell, uint32_t lobe) {
    float4 v = guide_lobes(p)[cell * (GUIDE_LOBES + 1u) + lobe];
    if (v.w > 0.0f) return v;
    const float k = 0.57735026919f;  // the tetrahedron's directions, normalized
    float sx = (lobe == 0 || lobe == 1) ? k : -k;
    float sy = (lobe == 0 || lobe == 2) ? k : -k;
    float sz = (lobe == 0 || lobe == 3) ? k : -k;
    return make_float4(sx, sy, sz, 1.0f);
}

__device__ __forceinline__ float4 guide_weights(const Params &p, uint32_t cell) {
    float4 w = guide_lobes(p)[cell * (GUIDE_LOBES + 1u) + GUIDE_LOBES];
    float total = w.x + w.y + w.z + w.w;
    return total > 0.0f ? w / total : make_float4(0.25f, 0.25f, 0.25f, 0.25f);
}

__device__ __forceinline__ float vmf_pdf(float4 lobe, V3 d) {
    float kappa = lobe.w;
    if (kappa < 1e-3f) return 1.0f / (4.0f * kPi);
    return kappa / (2.0f * kPi * (1.0f - expf(-2.0f * kappa))) * expf(kappa * (dot(v3(lobe), d) - 1.0f));
}

__device__ __forceinline__ V3 vmf_sample(float4 lobe, float ux, float uy) {
    float kappa = fmaxf(lobe.w, 1e-3f);
    float w = 1.0f + logf(ux + (1.0f - ux) * expf(-2.0f * kappa)) / kappa;
    float s = sqrtf(fmaxf(0.0f, 1.0f - w * w));
    float phi = 2.0f * kPi * uy;
    V3 t, b;
    basis(v3(lobe), t, b);
    return normalize(t * (s * cosf(phi)) + b * (s * sinf(phi)) + v3(lobe) * w);
}

__device__ __forceinline__ float guide_pdf(const Params &p, uint32_t cell, V3 d) {
    float4 w = guide_weights(p, cell);
    float pdf = 0.0f;
    for (uint32_t k = 0; k < GUIDE_LOBES; k++) pdf += lane(w, (int)k) * vmf_pdf(guide_lobe(p, cell, k), d);
    return pdf;
}

__device__ __forceinline__ V3 guide_sample(const Params &p, uint32_t cell, float ux, float uy, float uz) {
    float4 w = guide_weights(p, cell);
    uint32_t lobe = GUIDE_LOBES - 1u;
    float pick = uz;
    for (uint32_t k = 0; k < GUIDE_LOBES; k++) {
        if (pick < lane(w, (int)k)) {
            lobe = k;
            break;
        }
        pick -= lane(w, (int)k);
    }
    return vmf_sample(guide_lobe(p, cell, lobe), ux, uy);
}

// The learned guide's mixture of a path's current vertex as inference stores
// it: two 16-byte words a path slot (lobes 0-1, then 2-3, `pool` words
// apart), each lobe its mean's octahedral code and a (kappa, weight) pair of
// halves. Decoding renormalises the weights, so sampling and density read one
// distribution.
struct LearnedMix {
    uint4 lo, hi;
};

__device__ __forceinline__ uint32_t half_pair(float a, float b) {
    return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)a) |
           ((uint32_t)__builtin_bit_cast(uint16_t, (_Float16)b) << 16);
}
__device__ __forceinline__ float half_lo(uint32_t x) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(x & 0xffffu)); }
__device__ __forceinline__ float half_hi(uint32_t x) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(x >> 16)); }

__device__ __forceinline__ void store_learned_mix(const Params &p, uint32_t path, const V3 (&mu)[GUIDE_LOBES],
                                                  const float (&kappa)[GUIDE_LOBES], const float (&w)[GUIDE_LOBES]) {
    uint4 *m = p.guide_mix;
    m[path] = make_uint4(octahedral_encode(mu[0]), half_pair(kappa[0], w[0]), octahedral_encode(mu[1]),
                         half_pair(kappa[1], w[1]));
    m[(size_t)p.pool + path] = make_uint4(octahedral_encode(mu[2]), half_pair(kappa[2], w[2]),
                                          octahedral_encode(mu[3]), half_pair(kappa[3], w[3]));
}
__device__ __forceinline__ LearnedMix learned_mix(const Params &p, uint32_t path) {
    const uint4 *m = p.guide_mix;
    return LearnedMix{m[path], m[(size_t)p.pool + path]};
}
__device__ __forceinline__ uint32_t mix_word(const LearnedMix &m, uint32_t i) {
    const uint4 &q = i < 4u ? m.lo : m.hi;
    uint32_t j = i & 3u;
    return j == 0u ? q.x : j == 1u ? q.y : j == 2u ? q.z : q.w;
}
// The learned lobes' vMF density and sampling use the hardware's
// approximate exp, log, sine, cosine, reciprocal and reciprocal square root
// (relative error about 1e-6): sampling and density read the same lobes, so
// only their agreement matters, not a reference's rounding (the grid's lobes
// keep the exact functions the oracle checks).
__device__ __forceinline__ V3 mix_mean(const LearnedMix &m, uint32_t k) {
    uint32_t code = mix_word(m, 2u * k);
    float x = fmaxf(-1.0f, (float)(int16_t)(code & 0xffffu) * (1.0f / 32767.0f));
    float y = fmaxf(-1.0f, (float)(int16_t)(code >> 16) * (1.0f / 32767.0f));
    V3 n{x, y, 1.0f - fabsf(x) - fabsf(y)};
    if (n.z < 0.0f) {
        float nx = (1.0f - fabsf(n.y)) * (n.x >= 0.0f ? 1.0f : -1.0f);
        float ny = (1.0f - fabsf(n.x)) * (n.y >= 0.0f ? 1.0f : -1.0f);
        n.x = nx;
        n.y = ny;
    }
    return n * __builtin_amdgcn_rsqf(dot(n, n));
}
__device__ __forceinline__ float mix_kappa(const LearnedMix &m, uint32_t k) { return half_lo(mix_word(m, 2u * k + 1u)); }
__device__ __forceinline__ float4 mix_weights(const LearnedMix &m) {
    float4 w = make_float4(half_hi(m.lo.y), half_hi(m.lo.w), half_hi(m.hi.y), half_hi(m.hi.w));
    float total = w.x + w.y + w.z + w.w;
    return total > 0.0f ? w * __builtin_amdgcn_rcpf(total) : make_float4(0.25f, 0.25f, 0.25f, 0.25f);
}
__device__ __forceinline__ float learned_vmf_pdf(V3 mu, float kappa, V3 d) {
    if (kappa < 1e-3f) return 1.0f / (4.0f * kPi);
    return kappa * __builtin_amdgcn_rcpf(2.0f * kPi * (1.0f - __expf(-2.0f * kappa))) * __expf(kappa * (dot(mu, d) - 1.0f));
}
__device__ __forceinline__ float learned_pdf(const LearnedMix &m, V3 d) {
    float4 w = mix_weights(m);
    float pdf = 0.0f;
#pragma unroll
    for (uint32_t k = 0; k < GUIDE_LOBES; k++) pdf += lane(w, (int)k) * learned_vmf_pdf(mix_mean(m, k), mix_kappa(m, k), d);
    return pdf;
}
__device__ __forceinline__ V3 learned_sample(const LearnedMix &m, float ux, float uy, float uz) {
    float4 w = mix_weights(m);
    uint32_t lobe = GUIDE_LOBES - 1u;
    float pick = uz;
    for (uint32_t k = 0; k < GUIDE_LOBES; k++) {
        if (pick < lane(w, (int)k)) {
            lobe = k;
            break;
        }
        pick -= lane(w, (int)k);
    }
    V3 mu = mix_mean(m, lobe);
    float kappa = fmaxf(mix_kappa(m, lobe), 1e-3f);
    float c = 1.0f + __logf(ux + (1.0f - ux) * __expf(-2.0f * kappa)) * __builtin_amdgcn_rcpf(kappa);
    c = fminf(1.0f, fmaxf(-1.0f, c));
    float s = sqrtf(fmaxf(0.0f, 1.0f - c * c));
    float phi = 2.0f * kPi * uy;
    V3 t, b;
    basis(mu, t, b);
    V3 r = t * (s * __cosf(phi)) + b * (s * __sinf(phi)) + mu * c;
    return r * __builtin_amdgcn_rsqf(dot(r, r));
}

// guide.glsl:121-139: credit the lobe that explains d best.
__device__ __forceinline__ void guide_train(const Params &p, uint32_t cell, V3 d, float value, float pdf) {
    if (cell == kNoneU || !(value > 0.0f) || !(pdf > 0.0f) || isinf(value)) return;
    float4 w = guide_weights(p, cell);
    uint32_t best = 0;
    float best_score = -1.0f;
    for (uint32_t k = 0; k < GUIDE_LOBES; k++) {
        float score = (lane(w, (int)k) + 0.05f) * vmf_pdf(guide_lobe(p, cell, k), d);
        if (score > best_score) {
            best_score = score;
            best = k;
        }
    }
    float x = fminf(value / pdf, 1e4f);
    float *s = guide_stats(p) + (size_t)(cell * GUIDE_LOBES + best) * 5u;
    atomicAdd(&s[0], x);
    atomicAdd(&s[1], x * d.x);
    atomicAdd(&s[2], x * d.y);
    atomicAdd(&s[3], x * d.z);
    atomicAdd(&s[4], 1.0f);
}

}  // namespace rt
}  // namespace lr
