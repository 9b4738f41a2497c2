// Surface shading of the wavefront integrator (shaders/shade_surface.comp,
// light_paths.glsl): the camera-path and light-path shading and their
// kernel. Included by render_kernels.hip (the plain and grid-guided
// variants) and render_shade_learned.hip (the learned guide's).
#pragma once

#include "render_connect.h"
#include "render_program.h"

namespace lr {
namespace rt {

// ---------------------------------------------------------- surface shading

struct Pushes {
    bool closest = false;
    bool shadow = false;
};

template <bool SPECTRAL>
__device__ __forceinline__ V3 shading_normal(const Params &p, uint4 attributes, float4 hit, V3 geometric) {
    V3 n = normalize(octahedral_decode(attributes.x) * (1.0f - hit.z - hit.w) + octahedral_decode(attributes.y) * hit.z +
                     octahedral_decode(attributes.z) * hit.w);
    if (dot(n, geometric) < 0.0f) n = -n;
    return n;
}

__device__ __forceinline__ bool disperses(const Params &p, uint32_t m, bool spectral) {
    return spectral && material_at(p, m, 12).w != 0.0f;
}

template <bool SPECTRAL> __device__ void dispersive_hit(Spec &throughput, Spec &radiance, Spec &lambda) {
    if (lambda.y < 0.0f) return;
    throughput = make_float4(throughput.x * 4.0f, 0.0f, 0.0f, 0.0f);
    radiance = make_float4(radiance.x * 4.0f, 0.0f, 0.0f, 0.0f);
    lambda = make_float4(lambda.x, -lambda.y, -lambda.z, -lambda.w);
}

// The density of lobe-wise guided sampling: the BSDF's (bsdf_pdf) with
// `guided` of its cosine-sampled diffuse choice (p_diffuse cos / pi,
// guided <= p_diffuse) drawn from the guide instead.
__device__ __forceinline__ float lobewise_pdf(float bsdf_pdf, float guide, float cos_in, float guided) {
    return fmaxf(0.0f, bsdf_pdf - guided * fmaxf(cos_in, 0.0f) * (1.0f / kPi)) + guided * guide;
}

// shade_surface.comp:48-312. PROGRAMS: some material has a node-graph
// program (shader_program.glsl), run once a hit before the surface is read.
// LOBEWISE (the learned guide, render_shade_learned.hip): the guide draws in
// place of part of the diffuse lobe's cosine sampling and the glossy lobes
// keep their full rate. Otherwise the grid's one-sample mixture over the
// whole BSDF (shade_surface.comp:155-164, 243-271).
template <bool SPECTRAL, bool PROGRAMS, bool LOBEWISE = false>
__device__ Pushes shade_camera(const Params &p, const PassState &ps, uint32_t path) {
    State st(p);
    Pushes out;
    uint32_t pixel = st.u(SU_PIXEL, path);
    uint32_t bounces = st.u(SU_BOUNCE, path);
    uint32_t total = bounce_total(bounces);
    float4 hit = st.v(SV_HIT, path);
    V3 origin = v3(st.v(SV_ORIGIN, path));
    float4 ray = st.v(SV_DIRECTION, path);
    V3 rd = v3(ray);
    Spec throughput = st.v(SV_THROUGHPUT, path);
    Spec lambda = st.v(SV_LAMBDA, path);
    uint32_t what = st.u(SU_HIT, path);
    float4 vertex_row = st.v(SV_VERTEX, path);
    V3 vertex = v3(vertex_row);
    float4 mis = light_tracing(p) ? st.v(SV_MIS, path) : make_float4(0, 0, 0, 0);
    bool tracing = light_tracing(p);
    // The learned guide's record of the previous vertex holds the radiance
    // found before its sampled direction's light, counting that vertex's light
    // sample as arrived. Its shadow ray is resolved by now: the radiance here,
    // before this vertex adds any, counts it only if it arrived. (Counted as
    // arrived, a blocked light sample leaves the direction's light minus that
    // sample, mostly clipped to nothing where the lights are hidden.)
    if (LOBEWISE && ps.sample < p.train_samples && total >= 1u && total <= GUIDE_RING &&
        st.u(SU_GUIDE_CELL0 + total - 1u, path) != kNoneU) {
        float4 record = st.v(SV_GUIDE0 + total - 1u, path);
        record.z = spec_luminance<SPECTRAL>(p, st.v(SV_RADIANCE, path), lambda);
        st.set_v(SV_GUIDE0 + total - 1u, path, record);
    }

    if (what & kLightHit) {
        uint32_t light = what & ~kLightHit;
        float weight = 1.0f;
        if (total == 0u) {
            if (!light_visible(p, light)) return out;
        } else if (ray.w > 0.0f) {
            uint32_t emitter = light_emitter(p, light);
            float reach = distance(vertex, origin + rd * hit.x);
            float chance = (1.0f - K(p, K_TREE).z) * tree_pdf(p, emitter, vertex, octahedral_decode(st.u(SU_NORMAL, path)));
            float traced = tracing ? camera_hit_ratio(p, mis, emitter, light_cos(p, light, vertex, rd, reach), reach, ray.w) : 0.0f;
            weight = mis_weight(light_hit_pdf(p, light, vertex, rd, reach) * chance / ray.w, traced);
        }
        st.set_v(SV_RADIANCE, path,
                 st.v(SV_RADIANCE, path) + clamped<SPECTRAL>(p, throughput * light_radiance<SPECTRAL>(p, light, lambda) * weight, total));
        return out;
    }

    uint32_t tri = what;
    V3 p0, e1, e2;
    triangle_edges(p, tri, p0, e1, e2);
    V3 geometric = normalize(cross(e1, e2));
    bool inside = dot(geometric, rd) > 0.0f;
    if (inside) geometric = -geometric;
    uint4 attributes = p.shading[tri];
    V3 x = origin + rd * hit.x;
    float reach = distance(vertex, x);
    float density = (total == 0u ? mis.w : ray.w) * fabsf(dot(rd, geometric)) / (reach * reach);
    float back_reach = mis.z / (reach * reach);
    const uint32_t sample = slot_identity(ps, path).sample;
    float4 u_opacity = sample4(p.seed, pixel, sample, group_opacity(total));

    float opacity = clampf(material_at(p, attributes.w, 12).x, 0.0f, 1.0f);
    if (opacity < 1.0f &&
        fractf(u_opacity.x + (float)hash_u32(__float_as_uint(hit.x)) * (1.0f / 4294967296.0f)) >= opacity) {
        st.set_v(SV_ORIGIN, path, f4(offset_origin(x, -geometric), 0.0f));
        out.closest = true;
        return out;
    }

    if (disperses(p, attributes.w, SPECTRAL)) {
        Spec radiance = st.v(SV_RADIANCE, path);
        dispersive_hit<SPECTRAL>(throughput, radiance, lambda);
        st.set_v(SV_RADIANCE, path, radiance);
        st.set_v(SV_LAMBDA, path, lambda);
    }
    float hero = lambda.y < 0.0f ? lambda.x : 0.0f;

    if (material_at(p, attributes.w, 7).w > 0.0f) {
        float weight = 1.0f;
        uint32_t emitter = triangle_emitter(p, tri);
        if (total > 0u && ray.w > 0.0f && emitter != kNoneU) {
            float chance = (1.0f - K(p, K_TREE).z) * tree_pdf(p, emitter, vertex, octahedral_decode(st.u(SU_NORMAL, path)));
            float traced = tracing ? camera_hit_ratio(p, mis, emitter, fabsf(dot(rd, geometric)), reach, ray.w) : 0.0f;
            weight = mis_weight(triangle_hit_pdf(p, tri, rd, reach) * chance / ray.w, traced);
        }
        st.set_v(SV_RADIANCE, path,
                 st.v(SV_RADIANCE, path) +
                     clamped<SPECTRAL>(p, throughput * material_color<SPECTRAL>(p, attributes.w, 6u, lambda) * weight, total));
    }

    if (total >= (uint32_t)K(p, K_INFO).y) return out;

    V3 normal = shading_normal<SPECTRAL>(p, attributes, hit, geometric);
    // The material's node graph, once for this hit, on the outward normal; it
    // may bend the normal. A program that reads the view makes the light
    // tracer's density past the first vertex zero here (`viewed`).
    Program program;
    const Program *g = nullptr;
    bool viewed = false;
    if constexpr (PROGRAMS) {
        run_program(p, program, attributes.w, tri, x, inside ? -normal : normal, -rd,
                    triangle_uv(p, tri, hit.z, hit.w));
        g = &program;
        viewed = program.viewed;
        if (program.driven(9)) {
            normal = normalize(program.driven3(11));
            if (dot(normal, geometric) < 0.0f) normal = -normal;
        }
    }
    Surface surface = surface_at(normal, geometric, attributes.w, inside);
    V3 wo = to_local(surface, -rd);
    prepare_surface<SPECTRAL>(p, surface, wo, lambda, hero, g);

    V3 from = offset_origin(x, geometric);

    uint32_t cell = kNoneU;
    float guided = 0.0f;
    bool guide_on = guiding(p);
    if (guide_on && total < GUIDE_RING) {
        cell = p.learned ? path : guide_cell(p, x, geometric);
        bool suited = !surface.inside && surface.p_transmission <= 0.0f &&
                      (surface.p_diffuse > 0.05f || fminf(surface.alpha.x, surface.alpha.y) > 0.1f);
        bool ready = cell != kNoneU && (p.learned ? net_ready(p) : guide_count(p, cell) >= K(p, K_GUIDE).w);
        // The learned guide takes a share of the diffuse lobe's choice: its
        // mixtures model incident light times the cosine, the diffuse lobe's
        // ideal density, while a glossy lobe's own sampling follows its sharp
        // peak better (a mixture over the whole BSDF takes samples from that
        // peak and raises its variance). The grid takes a fixed share of
        // every BSDF sample (loading.lucb:281).
        // With light tracing, light paths connect to the camera from the
        // visible surfaces and estimate their diffuse light; the learned
        // guide then takes the second vertex alone (guiding the first as well
        // raised the glossy workload's error: 1.17e-3 against 1.01e-3 at 64
        // samples, 960x540; guiding the third as well left it the same,
        // 1.088e-3 against 1.085e-3, for 3% more time).
        if (suited && ready && (!LOBEWISE || !tracing || total == 1u))
            guided = K(p, K_GUIDE).z * (p.learned ? surface.p_diffuse : 1.0f);
    }
    // The learned mixture (32 bytes) is read once for every use below; at the
    // first vertex without light paths, the one inferred at the pixel's first
    // sample of the pass (its slot modulo the band).
    LearnedMix mix{};
    if constexpr (LOBEWISE)
        if (guided > 0.0f) mix = learned_mix(p, total == 0u && !tracing ? path % ps.band_count : path);

    float4 u_light = sample4(p.seed, pixel, sample, group_light(total));
    float p_distant = K(p, K_TREE).z;
    LightSample s;
    bool pushed_shadow = false;
    Spec emitted = spec(0.0f);
    float chance = 0.0f;
    bool sampled = false;
    uint32_t emitter = kNoneU, afar = kNoneU;
    if (u_light.x < p_distant) {
        float share;
        uint32_t light = pick_distant(p, u_light.x / p_distant, share);
        if (light != kNoneU && share > 0.0f && sample_light(p, light, x, u_light.y, u_light.z, s)) {
            afar = light;
            chance = p_distant * share;
            emitted = light_type(p, light) == LIGHT_DOME ? dome_radiance<SPECTRAL>(p, light, s.direction, lambda)
                                                         : light_radiance<SPECTRAL>(p, light, lambda);
            sampled = true;
        }
    } else {
        float tree_chance;
        if (tree_sample(p, (u_light.x - p_distant) / fmaxf(1e-7f, 1.0f - p_distant), from, normal, emitter, tree_chance)) {
            chance = (1.0f - p_distant) * tree_chance;
            uint32_t code = emitter_code(p, emitter);
            if (code & kEmitterTriangle) {
                sampled = sample_triangle<SPECTRAL>(p, code & ~kEmitterTriangle, x, u_light.y, u_light.z, lambda, s, emitted);
            } else if (sample_light(p, code, x, u_light.y, u_light.z, s)) {
                emitted = light_radiance<SPECTRAL>(p, code, lambda);
                sampled = true;
            }
        }
    }
    if (sampled && chance > 0.0f) {
        float side = dot(s.direction, geometric);
        V3 wi = to_local(surface, s.direction);
        if ((side > 0.0f) == (wi.z > 0.0f) && (side > 0.0f || surface.p_transmission > 0.0f)) {
            float bsdf_pdf;
            Spec f = evaluate_surface(surface, wi, wo, bsdf_pdf);
            if (guided > 0.0f && !tracing) bsdf_pdf = LOBEWISE ? lobewise_pdf(bsdf_pdf, learned_pdf(mix, s.direction), wi.z, guided) : guided * guide_pdf(p, cell, s.direction) + (1.0f - guided) * bsdf_pdf;
            if (bsdf_pdf > 0.0f) {
                float light_pdf = s.pdf * chance;
                float traced = 0.0f;
                if (tracing && !s.delta && emitter != kNoneU && density > 0.0f) {
                    float d2 = s.distance * s.distance;
                    float at_last = total == 0u ? 1.0f
                                    : (viewed ? 0.0f
                                              : light_sampler_pdf<SPECTRAL>(p, surface, s.direction, -rd, hero, g) *
                                                    back_reach / mis.y);
                    float at_vertex = emission_direction_pdf(p, emitter, s.cos_light) * fabsf(side) / d2 / density;
                    float at_light = emitter_area_pdf(p, emitter) / (bsdf_pdf * s.cos_light / d2);
                    traced = mis.x * at_last * at_vertex * at_light;
                }
                if (tracing && afar != kNoneU && emits_from_afar(p, afar) && density > 0.0f) {
                    float at_last = total == 0u ? 1.0f
                                    : (viewed ? 0.0f
                                              : light_sampler_pdf<SPECTRAL>(p, surface, s.direction, -rd, hero, g) *
                                                    back_reach / mis.y);
                    float at_vertex = afar_first_density(p, x, s.direction, fabsf(side)) / density;
                    float at_light = afar_emission_pdf(p, afar, s.direction) / bsdf_pdf;
                    traced = mis.x * at_last * at_vertex * at_light;
                }
                float weight = s.delta ? 1.0f : mis_weight(bsdf_pdf / light_pdf, traced * bsdf_pdf / light_pdf);
                Spec contribution = clamped<SPECTRAL>(p, throughput * f * fabsf(wi.z) * emitted * (weight / light_pdf), total);
                V3 target = x + s.direction * s.distance;
                float margin = 4e-4f * fmaxf(1.0f, fmaxf(fabsf(target.x), fmaxf(fabsf(target.y), fabsf(target.z))));
                float reach_s = s.distance >= kInf ? kInf : fmaxf(0.0f, s.distance - margin);
                st.set_v(SV_SHADOW_ORIGIN, path, f4(offset_origin(x, side > 0.0f ? geometric : -geometric), reach_s));
                st.set_v(SV_SHADOW_DIRECTION, path, f4(s.direction, 0.0f));
                st.set_v(SV_SHADOW_RADIANCE, path, contribution);
                out.shadow = true;
                pushed_shadow = true;
            }
        }
    }

    float4 u = sample4(p.seed, pixel, sample, group_bsdf(total));
    V3 wi;
    Spec weight;
    float pdf;
    uint32_t lobe;
    float unguided = 1.0f;
    float bsdf_only;
    if constexpr (LOBEWISE) {
        // The guide's direction or the BSDF's lobe choice (the diffuse lobe
        // keeping p_diffuse - guided of it), then one evaluation of the BSDF
        // for either: a wave's lanes take both choices, and the two
        // evaluations would run one after the other.
        bool from_guide = u_opacity.y < guided;
        if (from_guide) {
            wi = to_local(surface, learned_sample(mix, u_opacity.z, u_opacity.w, u.w));
            lobe = wi.z < 0.0f ? LOBE_TRANSMISSION : LOBE_DIFFUSE;
        } else if (!pick_direction(surface, wo, u.x, u.y, guided + u.w * (1.0f - guided), wi, lobe)) {
            return out;
        }
        float bsdf_pdf;
        Spec f = evaluate_surface(surface, wi, wo, bsdf_pdf);
        if (!from_guide && !(bsdf_pdf > 0.0f)) return out;
        bsdf_only = bsdf_pdf;
        pdf = guided > 0.0f ? lobewise_pdf(bsdf_pdf, learned_pdf(mix, to_world(surface, wi)), wi.z, guided) : bsdf_pdf;
        if (!(pdf > 0.0f) || (from_guide && spec_max<SPECTRAL>(f) <= 0.0f)) return out;
        unguided = from_guide ? pdf / fmaxf(bsdf_pdf, 1e-6f * pdf) : pdf / bsdf_pdf;
        weight = f * (fabsf(wi.z) / pdf);
    } else if (u_opacity.y < guided) {
        wi = to_local(surface, guide_sample(p, cell, u_opacity.z, u_opacity.w, u.w));
        float bsdf_pdf;
        Spec f = evaluate_surface(surface, wi, wo, bsdf_pdf);
        pdf = guided * guide_pdf(p, cell, to_world(surface, wi)) + (1.0f - guided) * bsdf_pdf;
        bsdf_only = bsdf_pdf;
        if (pdf <= 0.0f || spec_max<SPECTRAL>(f) <= 0.0f) return out;
        unguided = pdf / fmaxf(bsdf_pdf, 1e-6f * pdf);
        weight = f * (fabsf(wi.z) / pdf);
        lobe = wi.z < 0.0f ? LOBE_TRANSMISSION : (surface.p_diffuse > 0.5f ? LOBE_DIFFUSE : LOBE_GLOSSY);
    } else {
        if (!sample_surface(surface, wo, u.x, u.y, u.w, wi, weight, pdf, lobe)) return out;
        bsdf_only = pdf;
        if (guided > 0.0f) {
            float mixed = guided * guide_pdf(p, cell, to_world(surface, wi)) + (1.0f - guided) * pdf;
            unguided = mixed / pdf;
            weight = weight * (pdf / mixed);
            pdf = mixed;
        }
    }
    V3 direction = to_world(surface, wi);
    float side = dot(direction, geometric);
    if ((side > 0.0f) != (wi.z > 0.0f)) return out;
    bounces = next_bounce(p, bounces, lobe);
    if (bounces == kNoneU) return out;
    throughput = throughput * weight;
    float survival_scale = fminf(vertex_row.w * unguided, 1e6f);
    if (total + 1u >= 3u) {
        float survive = fminf(0.95f, spec_max<SPECTRAL>(throughput) * survival_scale);
        if (u.z >= survive) return out;
        throughput = throughput / survive;
    }
    st.set_u(SU_BOUNCE, path, bounces);
    st.set_u(SU_NORMAL, path, octahedral_encode(normal));
    // The learned guide reads records only after its training passes (the
    // grid after every pass).
    if (guide_on && total < GUIDE_RING && (!LOBEWISE || ps.sample < p.train_samples)) {
        if (cell == kNoneU) cell = p.learned ? path : guide_cell(p, x, geometric);
        st.set_u(SU_GUIDE_CELL0 + total, path, cell);
        if (p.learned)
            st.set_v(SV_GUIDE_POS0 + total, path, make_float4(x.x, x.y, x.z, __uint_as_float(octahedral_encode(geometric))));
        st.set_u(SU_GUIDE_DIR0 + total, path, octahedral_encode(direction));
        float found = spec_luminance<SPECTRAL>(p, st.v(SV_RADIANCE, path), lambda) +
                      (pushed_shadow ? spec_luminance<SPECTRAL>(p, st.v(SV_SHADOW_RADIANCE, path), lambda) : 0.0f);
        st.set_v(SV_GUIDE0 + total, path,
                 make_float4(pdf, spec_luminance<SPECTRAL>(p, throughput, lambda), found, fabsf(dot(direction, normal))));
        if (total + 1u < GUIDE_RING) st.set_u(SU_GUIDE_CELL0 + total + 1u, path, kNoneU);
    }
    V3 next = side > 0.0f ? from : offset_origin(x, -geometric);
    st.set_v(SV_ORIGIN, path, f4(next, 0.0f));
    st.set_v(SV_VERTEX, path, f4(next, survival_scale));
    if (tracing) {
        float at_last = total == 0u ? 1.0f
                        : (viewed ? 0.0f : light_sampler_pdf<SPECTRAL>(p, surface, direction, -rd, hero, g) * back_reach / mis.y);
        st.set_v(SV_MIS, path, make_float4(mis.x * at_last, density, fabsf(side), 0.0f));
        pdf = bsdf_only;
    }
    st.set_v(SV_DIRECTION, path, f4(direction, pdf));
    st.set_v(SV_THROUGHPUT, path, throughput);
    out.closest = true;
    return out;
}

// light_paths.glsl:166-180
template <bool SPECTRAL>
__device__ __forceinline__ float light_sample_ratio(const Params &p, uint32_t emitter, const Surface &s, V3 x, V3 x0,
                                                    V3 toward, float bsdf_pdf) {
    if (!(bsdf_pdf > 0.0f)) return 0.0f;
    V3 from = offset_origin(x, s.geometric);
    if (emitter & kAfarEmitter) {
        uint32_t light = emitter & ~kAfarEmitter;
        return light_select_pdf(p, light) * light_hit_pdf(p, light, from, toward, kInf) / bsdf_pdf;
    }
    V3 offset = x0 - from;
    float reach = length(offset);
    V3 d = offset / reach;
    float chance = (1.0f - K(p, K_TREE).z) * tree_pdf(p, emitter, from, s.normal);
    uint32_t code = emitter_code(p, emitter);
    float pdf = (code & kEmitterTriangle) ? triangle_hit_pdf(p, code & ~kEmitterTriangle, d, reach)
                                          : light_hit_pdf(p, code, from, d, reach);
    return chance * pdf / bsdf_pdf;
}

// light_paths.glsl:31-189
template <bool SPECTRAL, bool PROGRAMS>
__device__ Pushes shade_light(const Params &p, const PassState &ps, uint32_t path) {
    State st(p);
    Pushes out;
    uint32_t what = st.u(SU_HIT, path);
    if (what & kLightHit) return out;
    uint32_t emitter = st.u(SU_PIXEL, path);
    uint32_t bounces = st.u(SU_BOUNCE, path);
    uint32_t total = bounce_total(bounces);
    float4 hit = st.v(SV_HIT, path);
    V3 origin = v3(st.v(SV_ORIGIN, path));
    float4 ray = st.v(SV_DIRECTION, path);
    V3 rd = v3(ray);
    Spec throughput = st.v(SV_THROUGHPUT, path);
    Spec lambda = st.v(SV_LAMBDA, path);
    float4 vertex_row = st.v(SV_VERTEX, path);
    float4 mis = st.v(SV_MIS, path);
    const SlotIdentity who = slot_identity(ps, path);
    const uint32_t id = 0x80000000u | who.slot;

    uint32_t tri = what;
    V3 p0, e1, e2;
    triangle_edges(p, tri, p0, e1, e2);
    V3 geometric = normalize(cross(e1, e2));
    bool back = dot(geometric, rd) > 0.0f;
    if (back) geometric = -geometric;
    uint4 attributes = p.shading[tri];
    V3 x = origin + rd * hit.x;
    float4 u_opacity = sample4(p.seed, id, who.sample, group_opacity(total));

    float opacity = clampf(material_at(p, attributes.w, 12).x, 0.0f, 1.0f);
    if (opacity < 1.0f &&
        fractf(u_opacity.x + (float)hash_u32(__float_as_uint(hit.x)) * (1.0f / 4294967296.0f)) >= opacity) {
        st.set_v(SV_ORIGIN, path, f4(offset_origin(x, -geometric), 0.0f));
        out.closest = true;
        return out;
    }
    if (total >= (uint32_t)K(p, K_INFO).y) return out;
    if (disperses(p, attributes.w, SPECTRAL)) {
        Spec unused = spec(0.0f);
        dispersive_hit<SPECTRAL>(throughput, unused, lambda);
        st.set_v(SV_LAMBDA, path, lambda);
    }
    float hero = lambda.y < 0.0f ? lambda.x : 0.0f;

    V3 normal = shading_normal<SPECTRAL>(p, attributes, hit, geometric);
    uint32_t pixel = 0;
    V3 to_point;
    float camera_distance;
    float4 ul = sample4(p.seed, id, who.sample, group_light(total));
    V3 lens = lens_point(p, ul.x, ul.y);
    bool on_film = camera_pixel(p, x, lens, pixel, to_point, camera_distance);
    // The material's node graph as camera paths run it (a program reading the
    // view sees the point from the camera). Such a program has no surface for
    // directions the camera does not see from: the path connects here, if it
    // can, and goes no further.
    Program program;
    const Program *g = nullptr;
    bool viewed = false;
    if constexpr (PROGRAMS) {
        run_program(p, program, attributes.w, tri, x, back ? -normal : normal, on_film ? -to_point : -rd,
                    triangle_uv(p, tri, hit.z, hit.w));
        g = &program;
        viewed = program.viewed;
        if (viewed && !on_film) return out;
        if (program.driven(9)) {
            normal = normalize(program.driven3(11));
            if (dot(normal, geometric) < 0.0f) normal = -normal;
        }
    }

    Surface lit = surface_at(normal, geometric, attributes.w, back);
    V3 toward_light = -rd;
    read_material<SPECTRAL>(p, lit, hero, g);
    read_colors<SPECTRAL>(p, lit, lambda, g);
    V3 x_last = v3(vertex_row);
    float reach = distance(x_last, x);
    float cos_in = fabsf(dot(rd, geometric));
    float density = ray.w * cos_in / (reach * reach);
    float back_reach = mis.z / (reach * reach);
    if (total == 0u && (emitter & kAfarEmitter)) {
        density = afar_first_density(p, x, -rd, cos_in);
        back_reach = 1.0f;
        if (!(density > 0.0f)) return out;
    }
    float cos_shading_in = dot(toward_light, normal);
    if (cos_shading_in <= 0.0f || cos_in < 1e-6f) return out;
    float shading_in = cos_shading_in / cos_in;

    if (on_film) {
        V3 to_camera = -to_point;
        float side = dot(to_camera, geometric);
        Surface seen = side > 0.0f ? lit : turned(lit);
        V3 wo = to_local(seen, to_camera), wi = to_local(seen, toward_light);
        if ((dot(toward_light, seen.geometric) > 0.0f) == (wi.z > 0.0f) && wo.z > 0.0f) {
            if (side > 0.0f) orient_surface(p, seen, wo);
            else prepare_surface<SPECTRAL>(p, seen, wo, lambda, hero, g);
            float camera_bsdf_pdf;
            Spec f = evaluate_surface(seen, wi, wo, camera_bsdf_pdf);
            float camera_density = camera_pdf(p, to_point) * fabsf(side) / (camera_distance * camera_distance);
            Spec contribution = throughput * f * (shading_in * camera_density / light_paths_per_sample(p));
            if (spec_max<SPECTRAL>(contribution) > 0.0f) {
                float last = mis.y > 0.0f ? camera_bsdf_pdf * back_reach / mis.y : 0.0f;
                float ratio = mis.x * last * camera_density / density;
                float sample_ratio = total == 0u ? light_sample_ratio<SPECTRAL>(p, emitter, seen, x, x_last, toward_light, camera_bsdf_pdf)
                                                 : fmaxf(mis.w, 0.0f);
                float weight = mis_weight(ratio / light_paths_per_sample(p), ratio * sample_ratio / light_paths_per_sample(p));
                float margin = 4e-4f * fmaxf(1.0f, fmaxf(fabsf(x.x), fmaxf(fabsf(x.y), fabsf(x.z))));
                st.set_v(SV_SHADOW_ORIGIN, path, f4(offset_origin(x, seen.geometric), fmaxf(0.0f, camera_distance - margin)));
                st.set_v(SV_SHADOW_DIRECTION, path, f4(to_camera, 0.0f));
                st.set_v(SV_SHADOW_RADIANCE, path, f4(spec_to_film<SPECTRAL>(p, contribution * weight, lambda), 0.0f));
                st.set_u(SU_SPLAT, path, pixel);
                out.shadow = true;
            }
        }
    }

    if (viewed || total + 1u >= (uint32_t)K(p, K_INFO).y) return out;
    float4 u = sample4(p.seed, id, who.sample, group_bsdf(total));
    Surface choosing = lit;
    light_lobes(choosing);
    V3 wo_light = to_local(lit, toward_light);
    V3 wi_light;
    float light_pdf;
    uint32_t lobe;
    if (!sample_direction(choosing, wo_light, u.x, u.y, u.w, wi_light, light_pdf, lobe)) return out;
    V3 direction = to_world(lit, wi_light);
    float side = dot(direction, geometric);
    if ((side > 0.0f) != (wi_light.z > 0.0f)) return out;
    Surface seen = side > 0.0f ? lit : turned(lit);
    V3 wo = to_local(seen, direction), wi = to_local(seen, toward_light);
    if ((dot(toward_light, seen.geometric) > 0.0f) != (wi.z > 0.0f)) return out;
    if (side > 0.0f) orient_surface(p, seen, wo);
    else prepare_surface<SPECTRAL>(p, seen, wo, lambda, hero, g);
    float camera_bsdf_pdf;
    Spec f = evaluate_surface(seen, wi, wo, camera_bsdf_pdf);
    bounces = next_bounce(p, bounces, lobe);
    if (bounces == kNoneU) return out;
    throughput = throughput * f * (shading_in * fabsf(side) / light_pdf);
    float survival_scale = vertex_row.w;
    if (total + 1u >= 3u) {
        float survive = fminf(0.95f, spec_max<SPECTRAL>(throughput) * survival_scale);
        if (u.z >= survive) return out;
        throughput = throughput / survive;
    }
    if (!(spec_max<SPECTRAL>(throughput) > 0.0f)) return out;
    float last = mis.y > 0.0f ? camera_bsdf_pdf * back_reach / mis.y : 0.0f;
    float sample_ratio = total == 0u ? light_sample_ratio<SPECTRAL>(p, emitter, seen, x, x_last, toward_light, camera_bsdf_pdf)
                                     : mis.w;
    st.set_u(SU_BOUNCE, path, bounces);
    st.set_v(SV_ORIGIN, path, f4(offset_origin(x, side > 0.0f ? geometric : -geometric), 0.0f));
    st.set_v(SV_VERTEX, path, f4(x, survival_scale));
    st.set_v(SV_DIRECTION, path, f4(direction, light_pdf));
    st.set_v(SV_THROUGHPUT, path, throughput);
    st.set_v(SV_MIS, path, make_float4(mis.x * last, density, fabsf(side), sample_ratio));
    out.closest = true;
    return out;
}

// Camera paths' and light paths' hits are shaded by separate launches: each
// needs fewer registers than the two together, so more waves fit; capped at
// 168 registers (3 waves a SIMD, a few spilled values): measured faster than
// the uncapped 215 (2 waves) and than 128 (4 waves, more spills).
template <bool SPECTRAL, bool LIGHT, bool PROGRAMS, bool LOBEWISE = false>
__global__ __launch_bounds__(128, 3) void shade_kernel(const Params *plist, uint32_t n, uint32_t current) {
    __shared__ RowPlan plan;
    const uint32_t queue = LIGHT ? Q_LIGHT : Q_SURFACE;
    auto count_of = [](const Params &q) { return q.queues[LIGHT ? Q_LIGHT : Q_SURFACE]; };
    if (idle_block(plist, n, count_of)) return;
    uint32_t total = plan.make(plist, n, count_of);
    // consumed by this bounce's closest-hit and routing kernels
    if (!LIGHT && blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) plist[i].queues[Q_CLOSEST + current] = 0u;
    __shared__ StagedPushes<2, 128 * kStageIterations> staged;
    staged.init({Q_CLOSEST + 1u - current, Q_SHADOW});
    FOR_ROW_ITEMS(plan, n, total, g, r, at, active) {
        const Params &p = plist[r];
        const PassState &ps = *p.pass;
        Pushes pu;
        uint32_t path = 0;
        if (active) {
            path = queue_path(p, queue, at);
            if constexpr (LIGHT) pu = shade_light<SPECTRAL, PROGRAMS>(p, ps, path);
            else if constexpr (LOBEWISE && !PROGRAMS) {
                // Inlined, as the plain kernels' camera shading is (there the
                // compiler's own choice): a call's stack would triple the
                // spectral kernel's scratch.
                [[clang::always_inline]] pu = shade_camera<SPECTRAL, PROGRAMS, LOBEWISE>(p, ps, path);
            } else pu = shade_camera<SPECTRAL, PROGRAMS, LOBEWISE>(p, ps, path);
        }
        staged.push(0, path, pu.closest);
        staged.push(1, path, pu.shadow);
        staged.sync(p, row_ends(plan, n, total, g, r));
    }
}

}  // namespace rt
}  // namespace lr
