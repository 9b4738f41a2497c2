// Gaussian splats along a ray (shaders/splats.glsl), after 3D Gaussian Ray
// Tracing (Moenne-Loccoz et al. 2024); the records are packed by
// luce_render/splats.py after src/render/splats.lucb.
//
// A ray meets each Gaussian once, at its point of maximum response: there the
// splat covers alpha = min(0.99, opacity exp(-d^2/2)), d the ray's Mahalanobis
// distance from the centre, and shows its colour, the SH toward the ray's
// direction. Along a path segment the splats met in order of distance add
// radiance sum_i T_i alpha_i c_i and pass on T = prod (1 - alpha_i) of what
// lies beyond. Splats are a non-scattering layer that emits and absorbs:
// nothing samples them as lights, so camera paths gather their emission with
// weight 1, and every technique that carries light along a segment (BSDF
// rays, shadow rays, light paths and their camera connections) sees the same
// deterministic transmittance, which leaves the MIS weights untouched.
// Segments end by Russian roulette once less than 1% gets through.
//
// MI355X form: the k-buffer of the 32 nearest hits lives in registers (every
// index a constant after unrolling), the BVH walk's stack in LDS, lane
// interleaved; the splat stages are kernels of their own that run only for
// scenes with splats, so the closest-hit and shadow kernels keep their
// registers.
#pragma once

#include "render_bsdf.h"

namespace lr {
namespace rt {

constexpr uint32_t K_SPLAT_TREE = 26;  // z the splat count, w the alpha cull
constexpr uint32_t kSplatK = 32;       // hits gathered a round
constexpr float kSplatRoulette = 0.01f;
constexpr uint32_t kSplatStack = 48;

__device__ __forceinline__ bool splats_present(const Params &p) {
    return p.splat_count > 0u && K(p, K_SPLAT_TREE).z > 0.0f;
}

// `v` turned by the unit quaternion q (xyzw).
__device__ __forceinline__ V3 splat_turn(float4 q, V3 v) {
    V3 u{q.x, q.y, q.z};
    V3 t = cross(u, v) * 2.0f;
    return v + t * q.w + cross(u, t);
}

__device__ __forceinline__ float4 as_float4(uint4 r) {
    return make_float4(__uint_as_float(r.x), __uint_as_float(r.y), __uint_as_float(r.z), __uint_as_float(r.w));
}

// Splat s's alpha where the ray from o along unit d passes nearest its centre
// (in the splat's own metric) and the distance t there; 0 past 3 sigma or
// under the cull. In the splat's frame scaled to unit sigma the ray is
// o' + t d', and its nearest point's squared distance is |o' x d'|^2 / |d'|^2.
__device__ __forceinline__ float splat_alpha(const Params &p, uint32_t s, V3 o, V3 d, float &t) {
    float4 r0 = as_float4(p.splat_records[s * 4u]);
    float4 q = as_float4(p.splat_records[s * 4u + 1u]);
    float4 r2 = as_float4(p.splat_records[s * 4u + 2u]);
    V3 inverse{r2.x, r2.y, r2.z};
    V3 lo = splat_turn(q, o - V3{r0.x, r0.y, r0.z}) * inverse;
    V3 ld = splat_turn(q, d) * inverse;
    float dd = dot(ld, ld);
    t = -dot(lo, ld) / dd;
    V3 c = cross(lo, ld);
    float m2 = dot(c, c) / dd;
    if (!(m2 <= 9.0f)) return 0.0f;
    float alpha = fminf(0.99f, r0.w * expf(-0.5f * m2));
    return alpha >= K(p, K_SPLAT_TREE).w ? alpha : 0.0f;
}

__device__ __forceinline__ float splat_half(const Params &p, uint32_t first, uint32_t index) {
    uint32_t w = p.splat_words[first + (index >> 1)];
    uint16_t bits = (uint16_t)((index & 1u) ? (w >> 16) : (w & 0xFFFFu));
    return (float)__builtin_bit_cast(_Float16, bits);
}

__device__ __forceinline__ V3 splat_item(const Params &p, uint32_t first, uint32_t item) {
    return V3{splat_half(p, first, item * 3u), splat_half(p, first, item * 3u + 1u), splat_half(p, first, item * 3u + 2u)};
}

// The real SH bands 1..3 toward unit d in the SH frame, over `used` items
// from word `first` (the 3DGS constants and polynomials).
__device__ __forceinline__ V3 splat_bands(const Params &p, uint32_t first, uint32_t used, V3 d) {
    const float c1 = 0.4886025119029199f;
    float x = d.x, y = d.y, z = d.z;
    V3 c = splat_item(p, first, 0) * (-c1 * y) + splat_item(p, first, 1) * (c1 * z) - splat_item(p, first, 2) * (c1 * x);
    if (used < 8u) return c;
    float xx = x * x, yy = y * y, zz = z * z;
    c = c + splat_item(p, first, 3) * (1.0925484305920792f * x * y) + splat_item(p, first, 4) * (-1.0925484305920792f * y * z) +
        splat_item(p, first, 5) * (0.31539156525252005f * (2.0f * zz - xx - yy)) +
        splat_item(p, first, 6) * (-1.0925484305920792f * x * z) + splat_item(p, first, 7) * (0.5462742152960396f * (xx - yy));
    if (used < 15u) return c;
    c = c + splat_item(p, first, 8) * (-0.5900435899266435f * y * (3.0f * xx - yy)) +
        splat_item(p, first, 9) * (2.890611442640554f * x * y * z) +
        splat_item(p, first, 10) * (-0.4570457994644658f * y * (4.0f * zz - xx - yy)) +
        splat_item(p, first, 11) * (0.3731763325901154f * z * (2.0f * zz - 3.0f * xx - 3.0f * yy)) +
        splat_item(p, first, 12) * (-0.4570457994644658f * x * (4.0f * zz - xx - yy)) +
        splat_item(p, first, 13) * (1.445305721320277f * z * (xx - yy)) +
        splat_item(p, first, 14) * (-0.5900435899266435f * x * (xx - 3.0f * yy));
    return c;
}

__device__ __forceinline__ float splat_srgb_decode(float c) {
    return c <= 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f);
}

// Splat s's radiance along unit d: the DC (in the cloud's encoding) plus the
// bands, clamped at 0, decoded, Rec.709 to ACEScg, as a Spec.
template <bool SPECTRAL> __device__ __forceinline__ Spec splat_radiance(const Params &p, uint32_t s, V3 d, Spec lambda) {
    uint4 r2 = p.splat_records[s * 4u + 2u];
    uint4 r3 = p.splat_records[s * 4u + 3u];
    V3 e{__uint_as_float(r3.x), __uint_as_float(r3.y), __uint_as_float(r3.z)};
    uint32_t items = r3.w & 31u;
    if (items >= 3u) {
        uint32_t first = r2.w;
        float fx = splat_half(p, first, 0), fy = splat_half(p, first, 1), fz = splat_half(p, first, 2),
              fw = splat_half(p, first, 3);
        float norm = fmaxf(sqrtf(fx * fx + fy * fy + fz * fz + fw * fw), 1e-20f);
        float4 frame = make_float4(fx / norm, fy / norm, fz / norm, fw / norm);
        e = e + splat_bands(p, first + 2u, items, splat_turn(frame, d));
    }
    e = V3{fmaxf(e.x, 0.0f), fmaxf(e.y, 0.0f), fmaxf(e.z, 0.0f)};
    if (r3.w & 32u) e = V3{splat_srgb_decode(e.x), splat_srgb_decode(e.y), splat_srgb_decode(e.z)};
    V3 aces{0.613097f * e.x + 0.339523f * e.y + 0.047379f * e.z, 0.070194f * e.x + 0.916354f * e.y + 0.013452f * e.z,
            0.020616f * e.x + 0.109570f * e.y + 0.869815f * e.z};
    return spec_of_texel<SPECTRAL>(p, aces, lambda);
}

// A segment's seed: the path, the sample and the bounce, salted per use.
__device__ __forceinline__ uint32_t splat_seed(const Params &p, uint32_t sample, uint32_t path, uint32_t bounces,
                                               uint32_t salt) {
    return hash_u32(path ^ hash_u32(sample * 0x9e3779b9u + p.seed) ^ (bounces * 0x85ebca6bu) ^ salt);
}

__device__ __forceinline__ float splat_random(uint32_t seed, uint32_t round) {
    return (float)(hash_u32(seed ^ (round * 0x9e3779b9u)) >> 8) * (1.0f / 16777216.0f);
}

// Russian roulette on a transmittance under 1%: kept at 1% with probability
// T / 0.01, else 0.
__device__ __forceinline__ float splat_roulette(float kept, uint32_t seed, uint32_t round) {
    if (kept >= kSplatRoulette) return kept;
    return splat_random(seed, round) * kSplatRoulette < kept ? kSplatRoulette : 0.0f;
}

// The entry distance into a box within [t_lo, t_hi], or -1.
__device__ __forceinline__ float splat_box_entry(float4 a, float4 b, V3 o, V3 inv, float t_lo, float t_hi) {
    float x0 = (a.x - o.x) * inv.x, x1 = (b.x - o.x) * inv.x;
    float y0 = (a.y - o.y) * inv.y, y1 = (b.y - o.y) * inv.y;
    float z0 = (a.z - o.z) * inv.z, z1 = (b.z - o.z) * inv.z;
    float enter = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), t_lo));
    float leave = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), t_hi));
    return enter <= leave ? enter : -1.0f;
}

__device__ __forceinline__ V3 splat_inverse(V3 d) {
    auto safe = [](float v) { return fabsf(v) < 1e-20f ? (v < 0.0f ? -1e-20f : 1e-20f) : v; };
    return V3{1.0f / safe(d.x), 1.0f / safe(d.y), 1.0f / safe(d.z)};
}

// A splat tree node: (lo, first) and (hi, count).
__device__ __forceinline__ void splat_node(const Params &p, uint32_t i, float4 &a, float4 &b) {
    a = p.splat_nodes[i * 2u];
    b = p.splat_nodes[i * 2u + 1u];
}

// The k nearest hits of a round: distances ascending and the splats; empty
// slots hold t = kInf.
struct SplatHits {
    float t[kSplatK];
    uint32_t s[kSplatK];
    __device__ __forceinline__ void clear() {
#pragma unroll
        for (uint32_t i = 0; i < kSplatK; i++) {
            t[i] = kInf;
            s[i] = 0u;
        }
    }
    __device__ __forceinline__ bool full() const { return t[kSplatK - 1] < kInf; }
    // Insertion: each slot keeps the nearer of itself and the hit carried
    // down; the farthest falls off the end.
    __device__ __forceinline__ void keep(float ht, uint32_t hs) {
#pragma unroll
        for (uint32_t i = 0; i < kSplatK; i++) {
            bool nearer = ht < t[i];
            float kt = t[i];
            uint32_t ks = s[i];
            t[i] = nearer ? ht : kt;
            s[i] = nearer ? hs : ks;
            ht = nearer ? kt : ht;
            hs = nearer ? ks : hs;
        }
    }
};

// The kSplatK nearest splats whose maximum response lies in (t_lo, t_hi),
// near child first; nodes left on the stack are tested again against the
// cutoff, which falls once the buffer is full. `stack`: this lane's LDS
// column (stride `lanes`).
__device__ __forceinline__ void splat_gather(const Params &p, V3 o, V3 d, float t_lo, float t_hi, SplatHits &hits,
                                             uint32_t *stack, uint32_t lanes) {
    hits.clear();
    float cutoff = t_hi;
    V3 inv = splat_inverse(d);
    float4 a, b;
    splat_node(p, 0, a, b);
    if (a.x > b.x || splat_box_entry(a, b, o, inv, t_lo, cutoff) < 0.0f) return;
    uint32_t depth = 0, node = 0;
    while (true) {
        splat_node(p, node, a, b);
        uint32_t first = __float_as_uint(a.w), count = __float_as_uint(b.w);
        if (count > 0u) {
            for (uint32_t i = 0; i < count; i++) {
                float t;
                float alpha = splat_alpha(p, first + i, o, d, t);
                if (alpha <= 0.0f || !(t > t_lo && t < cutoff)) continue;
                hits.keep(t, first + i);
                if (hits.full()) cutoff = hits.t[kSplatK - 1];
            }
        } else {
            float4 la, lb, ra, rb;
            splat_node(p, first, la, lb);
            splat_node(p, first + 1u, ra, rb);
            float ea = splat_box_entry(la, lb, o, inv, t_lo, cutoff);
            float eb = splat_box_entry(ra, rb, o, inv, t_lo, cutoff);
            if (ea >= 0.0f && eb >= 0.0f) {
                bool left_first = ea <= eb;
                if (depth < kSplatStack) stack[(depth++) * lanes] = left_first ? first + 1u : first;
                node = left_first ? first : first + 1u;
                continue;
            }
            if (ea >= 0.0f) {
                node = first;
                continue;
            }
            if (eb >= 0.0f) {
                node = first + 1u;
                continue;
            }
        }
        bool found = false;
        while (depth > 0u && !found) {
            node = stack[(--depth) * lanes];
            splat_node(p, node, a, b);
            found = splat_box_entry(a, b, o, inv, t_lo, cutoff) >= 0.0f;
        }
        if (!found) break;
    }
}

// The splats along the ray up to t_hi, front to back, in rounds of the
// kSplatK nearest: the radiance they add (to be weighed by the path's
// throughput) and the fraction they pass on.
template <bool SPECTRAL>
__device__ __forceinline__ void splat_composite(const Params &p, V3 o, V3 d, float t_hi, Spec lambda, uint32_t seed,
                                                Spec &added, float &kept, uint32_t *stack, uint32_t lanes) {
    added = spec(0.0f);
    kept = 1.0f;
    float t_lo = 0.0f;
    for (uint32_t round = 0; round < 4096u; round++) {
        SplatHits hits;
        splat_gather(p, o, d, t_lo, t_hi, hits, stack, lanes);
#pragma unroll
        for (uint32_t i = 0; i < kSplatK; i++) {
            if (hits.t[i] < kInf) {
                float t;
                float alpha = splat_alpha(p, hits.s[i], o, d, t);
                added = added + splat_radiance<SPECTRAL>(p, hits.s[i], d, lambda) * (kept * alpha);
                kept *= 1.0f - alpha;
            }
        }
        if (!hits.full()) return;
        t_lo = hits.t[kSplatK - 1];
        kept = splat_roulette(kept, seed, round);
        if (kept == 0.0f) return;
    }
}

// The fraction the splats pass along the ray over (t_lo, t_hi): the product
// of their (1 - alpha), in any order.
__device__ __forceinline__ float splat_transmittance(const Params &p, V3 o, V3 d, float t_lo, float t_hi, uint32_t seed,
                                                     uint32_t *stack, uint32_t lanes) {
    float kept = 1.0f;
    uint32_t round = 0;
    V3 inv = splat_inverse(d);
    float4 a, b;
    splat_node(p, 0, a, b);
    if (a.x > b.x || splat_box_entry(a, b, o, inv, t_lo, t_hi) < 0.0f) return 1.0f;
    uint32_t depth = 0, node = 0;
    while (true) {
        splat_node(p, node, a, b);
        uint32_t first = __float_as_uint(a.w), count = __float_as_uint(b.w);
        if (count > 0u) {
            for (uint32_t i = 0; i < count; i++) {
                float t;
                float alpha = splat_alpha(p, first + i, o, d, t);
                if (alpha <= 0.0f || !(t > t_lo && t < t_hi)) continue;
                kept *= 1.0f - alpha;
                if (kept < kSplatRoulette) {
                    kept = splat_roulette(kept, seed, round++);
                    if (kept == 0.0f) return 0.0f;
                }
            }
        } else {
            float4 la, lb, ra, rb;
            splat_node(p, first, la, lb);
            splat_node(p, first + 1u, ra, rb);
            bool ha = splat_box_entry(la, lb, o, inv, t_lo, t_hi) >= 0.0f;
            bool hb = splat_box_entry(ra, rb, o, inv, t_lo, t_hi) >= 0.0f;
            if (ha && hb) {
                if (depth < kSplatStack) stack[(depth++) * lanes] = first + 1u;
                node = first;
                continue;
            }
            if (ha) {
                node = first;
                continue;
            }
            if (hb) {
                node = first + 1u;
                continue;
            }
        }
        if (depth == 0u) break;
        node = stack[(--depth) * lanes];
    }
    return kept;
}

}  // namespace rt
}  // namespace lr
