// Light-tracing connections (shaders/connect.glsl, common.glsl:113-143) and
// the indirect clamp, shared by the integrator's kernels.
#pragma once

#include "render_bsdf.h"
#include "render_guide.h"

namespace lr {
namespace rt {

// ---------------------------------------------------------- light tracing
// shaders/connect.glsl and common.glsl:113-143.

__device__ __forceinline__ bool light_tracing(const Params &p) { return K(p, K_CAMERA).x > 0.0f; }
__device__ __forceinline__ float light_paths_per_sample(const Params &p) { return K(p, K_CAMERA).x; }
__device__ __forceinline__ V3 frame_position(const Params &p) { return V3{p.frame[0], p.frame[1], p.frame[2]}; }
__device__ __forceinline__ V3 frame_corner(const Params &p) { return V3{p.frame[4], p.frame[5], p.frame[6]}; }
__device__ __forceinline__ V3 frame_du(const Params &p) { return V3{p.frame[8], p.frame[9], p.frame[10]}; }
__device__ __forceinline__ V3 frame_dv(const Params &p) { return V3{p.frame[12], p.frame[13], p.frame[14]}; }

__device__ __forceinline__ V3 lens_point(const Params &p, float ux, float uy) {
    float radius = K(p, K_CAMERA).z;
    if (radius <= 0.0f) return frame_position(p);
    float2 d = concentric(ux, uy);
    return frame_position(p) + normalize(frame_du(p)) * (d.x * radius) + normalize(frame_dv(p)) * (d.y * radius);
}

__device__ __forceinline__ float camera_pdf(const Params &p, V3 d) {
    float c = dot(d, v3(K(p, K_VIEW)));
    return c > 0.0f ? K(p, K_CAMERA).y / (c * c * c) : 0.0f;
}

__device__ __forceinline__ bool camera_pixel(const Params &p, V3 x, V3 lens, uint32_t &pixel, V3 &d, float &dist) {
    V3 offset = x - lens;
    dist = length(offset);
    d = offset / fmaxf(dist, 1e-20f);
    float c = dot(d, v3(K(p, K_VIEW)));
    if (c <= 1e-6f || dist <= 0.0f) return false;
    float focus = K(p, K_CAMERA).w;
    V3 film = (K(p, K_CAMERA).z > 0.0f ? (lens + d * (focus / c) - frame_position(p)) / focus : d / c) - frame_corner(p);
    V3 du = frame_du(p), dv = frame_dv(p);
    float fx = dot(film, du) / dot(du, du);
    float fy = dot(film, dv) / dot(dv, dv);
    if (!(fx >= 0.0f && fy >= 0.0f && fx < p.frame[11] && fy < p.frame[15])) return false;
    pixel = (uint32_t)fy * (uint32_t)p.frame[11] + (uint32_t)fx;
    return true;
}

__device__ __forceinline__ Surface turned(const Surface &s) {
    return surface_at(-s.normal, -s.geometric, s.material, !s.back);
}

template <bool SPECTRAL>
__device__ __forceinline__ float light_sampler_pdf(const Params &p, const Surface &s, V3 from, V3 to, float hero,
                                                   const Program *g = nullptr) {
    bool same_side = dot(from, s.geometric) >= 0.0f;
    Surface v = same_side ? s : turned(s);
    V3 wo = to_local(v, from), wi = to_local(v, to);
    if ((dot(to, v.geometric) > 0.0f) != (wi.z > 0.0f)) return 0.0f;
    if (!same_side) read_material<SPECTRAL>(p, v, hero, g);
    light_lobes(v);
    return surface_pdf(v, wi, wo);
}

__device__ __forceinline__ float emitter_power_pdf(const Params &p, uint32_t emitter) {
    uint32_t record = (uint32_t)K(p, K_EMITTERS).x + emitter * 4u;
    uint32_t trail = p.emitters[record + 1u], depth = p.emitters[record + 2u];
    uint32_t base = (uint32_t)K(p, K_TREE).y;
    uint32_t node = 0;
    float pdf = 1.0f;
    for (uint32_t level = 0; level < depth; level++) {
        uint32_t link = p.emitters[node * 2u];
        float left = fmaxf(p.lights[base + link * 3u].w, 0.0f);
        float right = fmaxf(p.lights[base + (link + 1u) * 3u].w, 0.0f);
        if (left + right <= 0.0f) return 0.0f;
        uint32_t bit = (trail >> level) & 1u;
        pdf *= (bit == 0u ? left : right) / (left + right);
        node = link + bit;
    }
    return pdf;
}

__device__ __forceinline__ bool emitter_by_power(const Params &p, float u, uint32_t &emitter, float &pdf) {
    pdf = 1.0f;
    emitter = 0;
    if (K(p, K_TREE).x < 1.0f) return false;
    uint32_t base = (uint32_t)K(p, K_TREE).y;
    uint32_t node = 0;
    for (uint32_t depth = 0; depth < 64u; depth++) {
        uint32_t link = p.emitters[node * 2u];
        if ((p.emitters[node * 2u + 1u] & kLeafNode) != 0u) {
            emitter = link;
            return pdf > 0.0f;
        }
        float left = fmaxf(p.lights[base + link * 3u].w, 0.0f);
        float right = fmaxf(p.lights[base + (link + 1u) * 3u].w, 0.0f);
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

__device__ __forceinline__ float emitter_area_pdf(const Params &p, uint32_t emitter) {
    float finite = 1.0f - K(p, K_EMIT).x;
    uint32_t code = emitter_code(p, emitter);
    float area;
    if (code & kEmitterTriangle) {
        V3 a, e1, e2;
        triangle_edges(p, code & ~kEmitterTriangle, a, e1, e2);
        area = 0.5f * length(cross(e1, e2));
    } else if (light_type(p, code) == LIGHT_SPHERE) {
        float r = light_at(p, code, 1).w;
        area = 4.0f * kPi * r * r;
    } else {
        area = light_at(p, code, 3).w;
    }
    return area > 0.0f ? finite * emitter_power_pdf(p, emitter) / area : 0.0f;
}

__device__ __forceinline__ float emission_direction_pdf(const Params &p, uint32_t emitter, float c) {
    bool two_sided = (emitter_code(p, emitter) & kEmitterTriangle) != 0u;
    return fabsf(c) / (two_sided ? 2.0f * kPi : kPi);
}

__device__ __forceinline__ float camera_hit_ratio(const Params &p, float4 mis, uint32_t emitter, float cos_light,
                                                  float dist, float bsdf_pdf) {
    if (emitter == kNoneU || mis.y <= 0.0f || bsdf_pdf <= 0.0f) return 0.0f;
    float d2 = dist * dist;
    float at_vertex = emission_direction_pdf(p, emitter, cos_light) * mis.z / d2 / mis.y;
    float at_light = emitter_area_pdf(p, emitter) / (bsdf_pdf * fabsf(cos_light) / d2);
    return mis.x * at_vertex * at_light;
}

template <bool SPECTRAL> __device__ __forceinline__ Spec clamped(const Params &p, Spec c, uint32_t total) {
    float clamp_value = K(p, K_INFO).z;
    if (total == 0u || clamp_value <= 0.0f) return c;
    float largest = spec_max<SPECTRAL>(c);
    return largest > clamp_value ? c * (clamp_value / largest) : c;
}

}  // namespace rt
}  // namespace lr
