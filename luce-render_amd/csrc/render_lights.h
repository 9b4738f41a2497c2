// UsdLux lights and the light tree on the device: sampling for next-event
// estimation, analytic hits for BSDF rays, and the pdfs MIS needs
// (shaders/lights.glsl:24-308, shaders/lighttree.glsl:18-137).
#pragma once

#include "render_common.h"

namespace lr {
namespace rt {

constexpr uint32_t LIGHT_STRIDE = 7;
constexpr uint32_t LIGHT_RECT = 0, LIGHT_DISK = 1, LIGHT_SPHERE = 2, LIGHT_DISTANT = 3, LIGHT_DOME = 4;
constexpr uint32_t kLeafNode = 0x80000000u;
constexpr uint32_t kEmitterTriangle = 0x80000000u;

__device__ __forceinline__ float4 light_at(const Params &p, uint32_t light, uint32_t row) {
    return p.lights[light * LIGHT_STRIDE + row];
}

// Element i of a float4 array that no kernel writes while this one runs, at a
// wave-uniform address (the analytic-light loops walk every light with a
// loop counter; a workgroup serves one request): read through the constant
// address space, so the wave issues one scalar load instead of a vector load
// of 64 equal addresses. The first active lane's pointer and index are used.
typedef float UniformFloat4 __attribute__((ext_vector_type(4)));
typedef const UniformFloat4 __attribute__((address_space(4))) *ConstFloat4;
__device__ __forceinline__ float4 uniform_at(const float4 *base, uint32_t i) {
    // readfirstlane returns a signed int: each half goes through uint32_t
    // before widening, or a low half with bit 31 set would sign-extend.
    uint64_t a = reinterpret_cast<uint64_t>(base);
    uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)a);
    uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    ConstFloat4 q = (ConstFloat4)((hi << 32) | lo);
    UniformFloat4 v = q[(uint32_t)__builtin_amdgcn_readfirstlane(i)];
    return make_float4(v.x, v.y, v.z, v.w);
}
// light_at for a wave-uniform light index.
__device__ __forceinline__ float4 uniform_light_at(const Params &p, uint32_t light, uint32_t row) {
    return uniform_at(p.lights, light * LIGHT_STRIDE + row);
}
__device__ __forceinline__ uint32_t light_type(const Params &p, uint32_t light) {
    return (uint32_t)light_at(p, light, 0).w;
}
__device__ __forceinline__ bool light_visible(const Params &p, uint32_t light) {
    return ((uint32_t)light_at(p, light, 5).w & 1u) != 0u;
}
__device__ __forceinline__ float light_select_pdf(const Params &p, uint32_t light) {
    return K(p, K_TREE).z * light_at(p, light, 6).y;
}
__device__ __forceinline__ uint32_t light_emitter(const Params &p, uint32_t light) {
    return (uint32_t)light_at(p, light, 6).w;
}
__device__ __forceinline__ uint32_t light_count(const Params &p) { return (uint32_t)K(p, K_INFO).x; }

template <bool SPECTRAL> __device__ __forceinline__ Spec light_radiance(const Params &p, uint32_t light, Spec lambda) {
    float4 rgb = light_at(p, light, 5);
    Spec r = spec_of<SPECTRAL>(light_at(p, light, 4), rgb, lambda);
    if constexpr (SPECTRAL) {
        if (((uint32_t)rgb.w & 2u) != 0u) {
            float kelvin = light_at(p, light, 6).x;
            r = make_float4(r.x * blackbody(lambda.x, kelvin), r.y * blackbody(lambda.y, kelvin),
                            r.z * blackbody(lambda.z, kelvin), r.w * blackbody(lambda.w, kelvin));
        }
    }
    return r;
}

// A distant or dome light by power; `pdf` its share among them.
__device__ __forceinline__ uint32_t pick_distant(const Params &p, float u, float &pdf) {
    uint32_t chosen = kNoneU;
    for (uint32_t k = 0; k < p.far_count; k++) {
        uint32_t l = p.far_lights[k];
        chosen = l;
        if (u < light_at(p, l, 6).z) break;
    }
    pdf = chosen == kNoneU ? 0.0f : light_at(p, chosen, 6).y;
    return chosen;
}

struct LightSample {
    V3 direction;
    float distance;
    float pdf;
    float cos_light;
    bool delta;
};

__device__ __forceinline__ V3 sample_cone(V3 axis, float cos_max, float ux, float uy) {
    float c = 1.0f - ux * (1.0f - cos_max);
    float s = sqrtf(fmaxf(0.0f, 1.0f - c * c));
    float phi = 2.0f * kPi * uy;
    V3 t, b;
    basis(axis, t, b);
    return normalize(t * (s * cosf(phi)) + b * (s * sinf(phi)) + axis * c);
}

__device__ __forceinline__ float sphere_hit(V3 center, float radius, V3 o, V3 d) {
    V3 oc = o - center;
    float b = dot(oc, d);
    float c = dot(oc, oc) - radius * radius;
    float h = b * b - c;
    if (h < 0.0f) return kInf;
    h = sqrtf(h);
    float t = -b - h;
    if (t > 0.0f) return t;
    t = -b + h;
    return t > 0.0f ? t : kInf;
}

// Dome lights (lights.glsl:89-175, environment.lucb): a lat-long image around
// the scene, its center down the light's -Z, its top +Y; a dome with no image
// (or not the one image K_DOME holds) shines its color evenly. Texels are
// constant, and sampling picks a row by the rows' cumulative weights, then a
// texel of it by the row's, so the pdf follows luminance x sin theta.
__device__ __forceinline__ bool dome_textured(const Params &p, uint32_t light) {
    return (float)light == K(p, K_DOME).x;
}

// The texel a world direction falls in, and sin theta there.
__device__ __forceinline__ uint32_t dome_texel(const Params &p, uint32_t light, V3 d, float &sin_theta) {
    V3 l{dot(d, v3(light_at(p, light, 1))), dot(d, v3(light_at(p, light, 2))), dot(d, v3(light_at(p, light, 3)))};
    uint32_t width = (uint32_t)K(p, K_DOME).z, height = (uint32_t)K(p, K_DOME).w;
    float u = 0.5f + atan2f(l.x, -l.z) * (0.5f / kPi);
    float v = acosf(clampf(l.y, -1.0f, 1.0f)) / kPi;
    sin_theta = sqrtf(fmaxf(0.0f, 1.0f - l.y * l.y));
    uint32_t x = min((uint32_t)(u * (float)width), width - 1u);
    uint32_t y = min((uint32_t)(v * (float)height), height - 1u);
    return y * width + x;
}

__device__ __forceinline__ float4 dome_texel_value(const Params &p, uint32_t texel) {
    return p.lights[(uint32_t)K(p, K_DOME).y + texel];
}

// The solid-angle pdf of sampling direction d (the light's selection chance
// not included): a texel's weight over the sum, spread over its solid angle.
__device__ __forceinline__ float dome_pdf(const Params &p, uint32_t light, V3 d) {
    if (!dome_textured(p, light)) return 1.0f / (4.0f * kPi);
    float sin_theta;
    uint32_t texel = dome_texel(p, light, d, sin_theta);
    if (sin_theta <= 0.0f) return 0.0f;
    float weight = dome_texel_value(p, texel).w;
    return weight / K(p, K_DOME_CDF).y * K(p, K_DOME).z * K(p, K_DOME).w / (2.0f * kPi * kPi * sin_theta);
}

template <bool SPECTRAL> __device__ __forceinline__ Spec dome_radiance(const Params &p, uint32_t light, V3 d, Spec lambda) {
    Spec radiance = light_radiance<SPECTRAL>(p, light, lambda);
    if (!dome_textured(p, light)) return radiance;
    float sin_theta;
    uint32_t texel = dome_texel(p, light, d, sin_theta);
    return radiance * spec_of_texel<SPECTRAL>(p, v3(dome_texel_value(p, texel)), lambda);
}

__device__ __forceinline__ float dome_cdf(const Params &p, uint32_t at) {
    return reinterpret_cast<const float *>(p.lights)[(uint32_t)K(p, K_DOME_CDF).x + at];
}

// The first of `count` cumulative values from `first` that passes u, and u's
// place within it (0..1).
__device__ __forceinline__ uint32_t dome_search(const Params &p, uint32_t first, uint32_t count, float u, float &within) {
    uint32_t low = 0, high = count - 1u;
    while (low < high) {
        uint32_t middle = (low + high) / 2u;
        if (dome_cdf(p, first + middle) > u) high = middle;
        else low = middle + 1u;
    }
    float below = low > 0u ? dome_cdf(p, first + low - 1u) : 0.0f;
    float above = dome_cdf(p, first + low);
    within = above > below ? clampf((u - below) / (above - below), 0.0f, 0.99999994f) : 0.5f;
    return low;
}

__device__ __forceinline__ bool sample_dome(const Params &p, uint32_t light, float ux, float uy, LightSample &s) {
    s.distance = kInf;
    s.delta = false;
    s.cos_light = 1.0f;
    V3 l;
    if (!dome_textured(p, light)) {
        float z = 1.0f - 2.0f * ux;
        float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        l = V3{r * cosf(2.0f * kPi * uy), z, r * sinf(2.0f * kPi * uy)};
    } else {
        uint32_t width = (uint32_t)K(p, K_DOME).z, height = (uint32_t)K(p, K_DOME).w;
        float fy, fx;
        uint32_t y = dome_search(p, width * height, height, uy, fy);
        uint32_t x = dome_search(p, y * width, width, ux, fx);
        float theta = kPi * ((float)y + fy) / (float)height;
        float phi = 2.0f * kPi * (((float)x + fx) / (float)width - 0.5f);
        l = V3{sinf(theta) * sinf(phi), cosf(theta), -sinf(theta) * cosf(phi)};
    }
    s.direction = normalize(v3(light_at(p, light, 1)) * l.x + v3(light_at(p, light, 2)) * l.y +
                            v3(light_at(p, light, 3)) * l.z);
    s.pdf = dome_pdf(p, light, s.direction);
    return s.pdf > 0.0f;
}

// lights.glsl:176-224
__device__ __forceinline__ bool sample_light(const Params &p, uint32_t light, V3 x, float ux, float uy,
                                             LightSample &s) {
    float4 r0 = light_at(p, light, 0), r1 = light_at(p, light, 1), r2 = light_at(p, light, 2),
           r3 = light_at(p, light, 3);
    uint32_t type = (uint32_t)r0.w;
    s.delta = false;
    s.cos_light = 1.0f;
    if (type == LIGHT_DOME) return sample_dome(p, light, ux, uy, s);
    if (type == LIGHT_RECT || type == LIGHT_DISK) {
        float ax, ay;
        if (type == LIGHT_RECT) {
            ax = 2.0f * ux - 1.0f;
            ay = 2.0f * uy - 1.0f;
        } else {
            float2 c = concentric(ux, uy);
            ax = c.x;
            ay = c.y;
        }
        V3 q = v3(r0) + v3(r1) * ax + v3(r2) * ay;
        V3 d = q - x;
        float d2 = dot(d, d);
        s.distance = sqrtf(d2);
        s.direction = d / s.distance;
        float cos_light = -dot(s.direction, v3(r3));
        if (cos_light <= 0.0f || r3.w <= 0.0f) return false;
        s.pdf = d2 / (cos_light * r3.w);
        s.cos_light = cos_light;
        return true;
    }
    if (type == LIGHT_SPHERE) {
        V3 to_center = v3(r0) - x;
        float d2 = dot(to_center, to_center);
        float r = r1.w;
        if (d2 <= r * r) return false;
        float cos_max = sqrtf(fmaxf(0.0f, 1.0f - r * r / d2));
        s.direction = sample_cone(to_center / sqrtf(d2), cos_max, ux, uy);
        s.distance = sphere_hit(v3(r0), r, x, s.direction);
        if (s.distance >= kInf) return false;
        s.pdf = 1.0f / (2.0f * kPi * (1.0f - cos_max));
        s.cos_light = fmaxf(0.0f, -dot(s.direction, normalize(x + s.direction * s.distance - v3(r0))));
        return true;
    }
    V3 axis = -v3(r3);
    float cos_half = r2.w;
    s.distance = kInf;
    if (cos_half >= 1.0f - 1e-7f) {
        s.direction = axis;
        s.pdf = 1.0f;
        s.delta = true;
        return true;
    }
    s.direction = sample_cone(axis, cos_half, ux, uy);
    s.pdf = 1.0f / (2.0f * kPi * (1.0f - cos_half));
    return true;
}

__device__ __forceinline__ float light_cos(const Params &p, uint32_t light, V3 o, V3 d, float t) {
    float4 r0 = light_at(p, light, 0);
    if ((uint32_t)r0.w == LIGHT_SPHERE) return fmaxf(0.0f, -dot(d, normalize(o + d * t - v3(r0))));
    return fmaxf(0.0f, -dot(d, v3(light_at(p, light, 3))));
}

// lights.glsl:239-259: the nearer hit on the emitting side, or kInf.
// UNIFORM: `light` is the same in every active lane (the loops below).
template <bool UNIFORM = false>
__device__ __forceinline__ float intersect_light(const Params &p, uint32_t light, V3 o, V3 d, float t_max,
                                                 bool light_flow) {
    auto at = [&](uint32_t row) { return UNIFORM ? uniform_light_at(p, light, row) : light_at(p, light, row); };
    float4 r0 = at(0);
    uint32_t type = (uint32_t)r0.w;
    if (type == LIGHT_SPHERE) {
        float t = sphere_hit(v3(r0), at(1).w, o, d);
        return t < t_max ? t : kInf;
    }
    if (type >= LIGHT_DISTANT) return kInf;
    V3 u = v3(at(1)), v = v3(at(2)), n = v3(at(3));
    float facing = dot(d, n);
    if (light_flow ? facing <= 0.0f : facing >= 0.0f) return kInf;
    float t = dot(v3(r0) - o, n) / facing;
    if (t <= 0.0f || t >= t_max) return kInf;
    V3 q = o + d * t - v3(r0);
    float a = dot(q, u) / dot(u, u);
    float b = dot(q, v) / dot(v, v);
    bool inside = type == LIGHT_RECT ? (fabsf(a) <= 1.0f && fabsf(b) <= 1.0f) : (a * a + b * b <= 1.0f);
    return inside ? t : kInf;
}

// Culling for the analytic lights a ray may meet: a tree over the lights'
// boxes (built on the host: light_tree, their centres' Morton order as the
// leaves of a complete binary tree), walked by each lane for its own ray. A
// node whose box the segment enters has its two children's boxes read
// together (one 64-byte read); the walk goes on into the first child it
// enters, and a bit per level (`trail`) marks a second child still to visit,
// found again by shifting the node index (an ancestor of node k at d levels
// up is k >> d, a left child is even). A wave-uniform loop over groups of
// lights would enter a group when any of 64 incoherent rays may meet it, so
// every wave would test nearly every light; here a lane tests the lights its
// own segment passes. Boxes are grown well past rounding, so a light
// intersect_light would return is never culled.
template <class Limit, class Visit>
__device__ __forceinline__ void walk_lights(const Params &p, V3 o, V3 d, Limit t_max, Visit visit) {
    const uint32_t count = p.light_nodes;
    if (count == 0u) return;
    const LightNode *nodes = p.light_tree;
    const uint32_t leaves = count / 2u;
    auto inverse = [](float x) { return 1.0f / (fabsf(x) > 1e-20f ? x : copysignf(1e-20f, x)); };
    V3 inv{inverse(d.x), inverse(d.y), inverse(d.z)};
    auto enters = [&](float4 lo, float4 hi) {
        float tx0 = (lo.x - o.x) * inv.x, tx1 = (hi.x - o.x) * inv.x;
        float ty0 = (lo.y - o.y) * inv.y, ty1 = (hi.y - o.y) * inv.y;
        float tz0 = (lo.z - o.z) * inv.z, tz1 = (hi.z - o.z) * inv.z;
        float enter = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
        float leave = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
        return enter <= leave && enter <= t_max();
    };
    {
        const float4 *root = reinterpret_cast<const float4 *>(nodes + 1);
        if (!enters(root[0], root[1])) return;
    }
    uint32_t k = 1u, depth = 0u, trail = 0u;
    for (;;) {
        if (k >= leaves) {
            uint32_t light = nodes[k].light;
            if (light != kNoneU && visit(light)) return;
        } else {
            const float4 *pair = reinterpret_cast<const float4 *>(nodes + 2u * k);
            bool left = enters(pair[0], pair[1]), right = enters(pair[2], pair[3]);
            if (left || right) {
                if (left && right) trail |= 1u << depth;
                k = left ? 2u * k : 2u * k + 1u;
                depth++;
                continue;
            }
        }
        // the deepest second child still to visit, if any
        if (trail == 0u) return;
        uint32_t level = 31u - (uint32_t)__builtin_clz(trail);
        trail &= ~(1u << level);
        k = (k >> (depth - level - 1u)) | 1u;
        depth = level + 1u;
    }
}

// The nearest analytic light before `t` along the ray (lights.glsl:239-259
// over every light, intersect_closest.comp:58-66): updates t and what. The
// reference's loop goes in index order with a strict `<`, so a triangle wins
// a tie with a light and the lowest index a tie between lights; the walk
// meets lights in tree order, so a light at exactly the distance of a
// higher-indexed light found before it is also taken (intersect_light
// searches one ulp past t while a higher index holds it).
__device__ __forceinline__ void nearest_light(const Params &p, V3 o, V3 d, bool camera_ray, bool light_flow, float &t,
                                              uint32_t &what) {
    walk_lights(p, o, d, [&] { return t; }, [&](uint32_t l) {
        if (camera_ray && !light_visible(p, l)) return false;
        // a light holds the hit (a miss, kNoneU, has the light bit too)
        bool later = what != kNoneU && (what & kLightHit) && (what & ~kLightHit) > l;
        float tl = intersect_light(p, l, o, d, later ? nextafterf(t, kInf) : t, light_flow);
        if (tl < t || (later && tl == t)) {
            t = tl;
            what = kLightHit | l;
        }
        return false;
    });
}

// Whether a light path's connection to the camera passes through a visible
// light before t_max (intersect_shadow.comp:55-60).
__device__ __forceinline__ bool light_in_the_way(const Params &p, V3 o, V3 d, float t_max) {
    bool blocked = false;
    walk_lights(p, o, d, [&] { return t_max; }, [&](uint32_t l) {
        blocked = light_visible(p, l) && intersect_light(p, l, o, d, t_max, true) < t_max;
        return blocked;
    });
    return blocked;
}

__device__ __forceinline__ float light_hit_pdf(const Params &p, uint32_t light, V3 o, V3 d, float t) {
    float4 r0 = light_at(p, light, 0);
    uint32_t type = (uint32_t)r0.w;
    if (type == LIGHT_SPHERE) {
        V3 to_center = v3(r0) - o;
        float r = light_at(p, light, 1).w;
        float cos_max = sqrtf(fmaxf(0.0f, 1.0f - r * r / dot(to_center, to_center)));
        return 1.0f / (2.0f * kPi * fmaxf(1e-12f, 1.0f - cos_max));
    }
    if (type == LIGHT_DISTANT) return 1.0f / (2.0f * kPi * fmaxf(1e-12f, 1.0f - light_at(p, light, 2).w));
    if (type == LIGHT_DOME) return dome_pdf(p, light, d);
    float4 r3 = light_at(p, light, 3);
    float cos_light = fmaxf(1e-12f, -dot(d, v3(r3)));
    return t * t / (cos_light * r3.w);
}

// Light paths from afar (lights.glsl:290-308).
__device__ __forceinline__ bool emits_from_afar(const Params &p, uint32_t light) {
    uint32_t type = light_type(p, light);
    return type == LIGHT_DOME || (type == LIGHT_DISTANT && light_at(p, light, 2).w < 1.0f - 1e-7f);
}
__device__ __forceinline__ float afar_emission_pdf(const Params &p, uint32_t light, V3 w) {
    return K(p, K_EMIT).x * light_at(p, light, 6).y * light_hit_pdf(p, light, V3{0, 0, 0}, w, kInf);
}
__device__ __forceinline__ float afar_first_density(const Params &p, V3 x, V3 w, float c) {
    float radius = K(p, K_EMIT).y;
    V3 offset = x - v3(K(p, K_EMIT_CENTER));
    V3 across = offset - w * dot(offset, w);
    if (dot(across, across) > radius * radius) return 0.0f;
    return c / (kPi * radius * radius);
}

// ---------------------------------------------------------------- light tree

__device__ __forceinline__ float safe_sqrt(float x) { return sqrtf(fmaxf(0.0f, x)); }

// pbrt-v4's importance of tree node `node` for point x with normal n
// (lighttree.glsl:18-47).
__device__ __forceinline__ float tree_importance(const Params &p, uint32_t node, V3 x, V3 n) {
    uint32_t base = (uint32_t)K(p, K_TREE).y + node * 3u;
    float4 a = p.lights[base], b = p.lights[base + 1], c = p.lights[base + 2];
    if (a.w <= 0.0f) return 0.0f;
    V3 center = (v3(a) + v3(b)) * 0.5f;
    V3 offset = x - center;
    float radius = 0.5f * length(v3(b) - v3(a));
    float d2o = dot(offset, offset);
    float distance2 = fmaxf(d2o, radius);
    V3 wi = d2o > 0.0f ? normalize(offset) : v3(c);
    float cos_w = dot(v3(c), wi);
    float sin_w = safe_sqrt(1.0f - cos_w * cos_w);
    float cos_b = d2o < radius * radius ? -1.0f : safe_sqrt(1.0f - radius * radius / d2o);
    float sin_b = safe_sqrt(1.0f - cos_b * cos_b);
    float cos_o = b.w;
    float sin_o = safe_sqrt(1.0f - cos_o * cos_o);
    float cos_x = cos_w > cos_o ? 1.0f : cos_w * cos_o + sin_w * sin_o;
    float sin_x = cos_w > cos_o ? 0.0f : sin_w * cos_o - cos_w * sin_o;
    float cos_p = cos_x > cos_b ? 1.0f : cos_x * cos_b + sin_x * sin_b;
    if (cos_p <= c.w) return 0.0f;
    float importance = a.w * cos_p / distance2;
    float cos_i = fabsf(dot(wi, n));
    float sin_i = safe_sqrt(1.0f - cos_i * cos_i);
    float cos_pi = cos_i > cos_b ? 1.0f : cos_i * cos_b + sin_i * sin_b;
    return fmaxf(importance * cos_pi, 0.0f);
}

__device__ __forceinline__ bool tree_sample(const Params &p, float u, V3 x, V3 n, uint32_t &emitter, float &pdf) {
    pdf = 1.0f;
    emitter = 0;
    if (K(p, K_TREE).x < 1.0f) return false;
    uint32_t node = 0;
    for (uint32_t depth = 0; depth < 64u; depth++) {
        uint32_t link = p.emitters[node * 2u];
        if ((p.emitters[node * 2u + 1u] & kLeafNode) != 0u) {
            emitter = link;
            return pdf > 0.0f;
        }
        float left = tree_importance(p, link, x, n);
        float right = tree_importance(p, link + 1u, x, n);
        if (left + right <= 0.0f) return false;
        float chance = left / (left + right);
        if (u < chance) {
            u = fminf(u / chance, 0.99999994f);
            pdf *= chance;
            node = link;
        } else {
            u = fminf((u - chance) / (1.0f - chance), 0.99999994f);
            pdf *= 1.0f - chance;
            node = link + 1u;
        }
    }
    return false;
}

__device__ __forceinline__ float tree_pdf(const Params &p, uint32_t emitter, V3 x, V3 n) {
    uint32_t record = (uint32_t)K(p, K_EMITTERS).x + emitter * 4u;
    uint32_t trail = p.emitters[record + 1u], depth = p.emitters[record + 2u];
    uint32_t node = 0;
    float pdf = 1.0f;
    for (uint32_t level = 0; level < depth; level++) {
        uint32_t link = p.emitters[node * 2u];
        float left = tree_importance(p, link, x, n);
        float right = tree_importance(p, link + 1u, x, n);
        if (left + right <= 0.0f) return 0.0f;
        uint32_t bit = (trail >> level) & 1u;
        pdf *= bit == 0u ? left / (left + right) : right / (left + right);
        node = link + bit;
    }
    return pdf;
}

__device__ __forceinline__ uint32_t emitter_code(const Params &p, uint32_t emitter) {
    return p.emitters[(uint32_t)K(p, K_EMITTERS).x + emitter * 4u];
}
__device__ __forceinline__ uint32_t triangle_emitter(const Params &p, uint32_t triangle) {
    return p.emitters[(uint32_t)K(p, K_EMITTERS).y + triangle];
}

// A leaf-order triangle's corners from the repacked triangle array (p0, e1, e2
// with e1 = p1 - p0, e2 = p2 - p0 as f32: the same edges shading computes).
__device__ __forceinline__ void triangle_edges(const Params &p, uint32_t tri, V3 &p0, V3 &e1, V3 &e2) {
    float4 a, b, c;
    load_triangle(p.scene.tris, tri, a, b, c);
    p0 = V3{a.x, a.y, a.z};
    e1 = V3{a.w, b.x, b.y};
    e2 = V3{b.z, b.w, c.x};
}

__device__ __forceinline__ float4 material_at(const Params &p, uint32_t m, uint32_t row) {
    return p.materials[m * 16u + row];
}
template <bool SPECTRAL> __device__ __forceinline__ Spec material_color(const Params &p, uint32_t m, uint32_t row,
                                                                       Spec lambda) {
    return spec_of<SPECTRAL>(material_at(p, m, row), material_at(p, m, row + 1u), lambda);
}

// An emissive triangle sampled uniformly by area (lighttree.glsl:107-129).
template <bool SPECTRAL>
__device__ __forceinline__ bool sample_triangle(const Params &p, uint32_t tri, V3 x, float ux, float uy, Spec lambda,
                                                LightSample &s, Spec &radiance) {
    V3 a, e1, e2;
    triangle_edges(p, tri, a, e1, e2);
    V3 b = a + e1, c = a + e2;
    float r = sqrtf(ux);
    V3 q = a * (1.0f - r) + b * (r * (1.0f - uy)) + c * (r * uy);
    V3 cross_ab = cross(e1, e2);
    float area = 0.5f * length(cross_ab);
    V3 d = q - x;
    float d2 = dot(d, d);
    if (area <= 0.0f || d2 <= 0.0f) return false;
    s.distance = sqrtf(d2);
    s.direction = d / s.distance;
    float cos_light = fabsf(dot(s.direction, cross_ab)) / (2.0f * area);
    if (cos_light <= 0.0f) return false;
    s.pdf = d2 / (cos_light * area);
    s.cos_light = cos_light;
    s.delta = false;
    radiance = material_color<SPECTRAL>(p, p.shading[tri].w, 6u, lambda);
    return true;
}

__device__ __forceinline__ float triangle_hit_pdf(const Params &p, uint32_t tri, V3 d, float t) {
    V3 a, e1, e2;
    triangle_edges(p, tri, a, e1, e2);
    V3 cross_ab = cross(e1, e2);
    float area = 0.5f * length(cross_ab);
    float cos_light = fabsf(dot(d, cross_ab)) / fmaxf(1e-20f, 2.0f * area);
    return t * t / fmaxf(1e-20f, cos_light * area);
}

}  // namespace rt
}  // namespace lr
