// GGX microfacets and the OpenPBR Surface lobes on the device
// (shaders/microfacet.glsl:9-84, shaders/openpbr.glsl:100-556): EON diffuse,
// GGX specular with Kulla–Conty compensation from albedo tables, F82-tint
// metal, a dielectric coat, rough dielectric transmission. Lobes are chosen
// by wavelength-free weights from the RGB parameters.
#pragma once

#include "render_lights.h"

namespace lr {
namespace rt {

// ------------------------------------------------------------------- GGX

__device__ __forceinline__ float2 ggx_alpha(float roughness, float anisotropy) {
    float r2 = roughness * roughness;
    float t = r2 * sqrtf(2.0f / (1.0f + (1.0f - anisotropy) * (1.0f - anisotropy)));
    return make_float2(fmaxf(t, 1e-4f), fmaxf((1.0f - anisotropy) * t, 1e-4f));
}
__device__ __forceinline__ float ggx_d(V3 m, float2 a) {
    if (m.z <= 0.0f) return 0.0f;
    V3 s{m.x / a.x, m.y / a.y, m.z};
    float k = dot(s, s);
    return 1.0f / (kPi * a.x * a.y * k * k);
}
__device__ __forceinline__ float ggx_lambda(V3 v, float2 a) {
    float z2 = v.z * v.z;
    if (z2 <= 0.0f) return kInf;
    float a2 = (a.x * a.x * v.x * v.x + a.y * a.y * v.y * v.y) / z2;
    return 0.5f * (-1.0f + sqrtf(1.0f + a2));
}
__device__ __forceinline__ float ggx_g1(V3 v, float2 a) { return 1.0f / (1.0f + ggx_lambda(v, a)); }
__device__ __forceinline__ float ggx_g2(V3 wi, V3 wo, float2 a) {
    return 1.0f / (1.0f + ggx_lambda(wi, a) + ggx_lambda(wo, a));
}
// Visible normals by spherical caps (Dupuy and Benyoub 2023).
__device__ __forceinline__ V3 ggx_sample(V3 wo, float2 a, float ux, float uy) {
    V3 v = normalize(V3{wo.x * a.x, wo.y * a.y, wo.z});
    float phi = 2.0f * kPi * ux;
    float z = (1.0f - uy) * (1.0f + v.z) - v.z;
    float s = sqrtf(clampf(1.0f - z * z, 0.0f, 1.0f));
    V3 h = V3{s * cosf(phi), s * sinf(phi), z} + v;
    return normalize(V3{h.x * a.x, h.y * a.y, fmaxf(h.z, 1e-6f)});
}
__device__ __forceinline__ float ggx_reflection_pdf(V3 wo, V3 h, float2 a) {
    return ggx_g1(wo, a) * ggx_d(h, a) / (4.0f * wo.z);
}
__device__ __forceinline__ float fresnel_dielectric(float cos_i, float eta) {
    float c = fabsf(cos_i);
    float g2 = eta * eta - 1.0f + c * c;
    if (g2 <= 0.0f) return 1.0f;
    float g = sqrtf(g2);
    float a = (g - c) / (g + c);
    float b = (c * (g + c) - 1.0f) / (c * (g - c) + 1.0f);
    return 0.5f * a * a * (1.0f + b * b);
}
__device__ __forceinline__ float fresnel_dielectric_average(float eta) {
    return logf((10893.0f * eta - 1438.2f) / (-774.4f * eta * eta + 10212.0f * eta + 1.0f));
}

// Albedo tables (microfacet.glsl:65-84): E_ggx(mu, r), its average over mu,
// and E_dielectric(mu, r, z) with z = sqrt(|eta - 1| / (eta + 1)).
constexpr uint32_t TABLE_MU = 32, TABLE_ROUGHNESS = 32, TABLE_ETA = 16;
constexpr float TABLE_Z_MAX = 0.8f, TABLE_MU_MIN = 0.02f;
constexpr uint32_t TABLE_GGX = 0, TABLE_GGX_AVERAGE = TABLE_MU * TABLE_ROUGHNESS,
                   TABLE_DIELECTRIC = TABLE_GGX_AVERAGE + TABLE_ROUGHNESS,
                   TABLE_SIZE = TABLE_DIELECTRIC + TABLE_MU * TABLE_ROUGHNESS * TABLE_ETA;

__device__ __forceinline__ float eta_of_z(float z) { return (1.0f + z * z) / (1.0f - z * z); }
__device__ __forceinline__ float z_of_eta(float eta) { return sqrtf(fabsf(eta - 1.0f) / (eta + 1.0f)); }
__device__ __forceinline__ float table_mu(uint32_t i) { return mixf(TABLE_MU_MIN, 1.0f, (float)i / (float)(TABLE_MU - 1)); }
__device__ __forceinline__ float table_roughness(uint32_t i) { return (float)i / (float)(TABLE_ROUGHNESS - 1); }
__device__ __forceinline__ float table_z(uint32_t i) { return TABLE_Z_MAX * (float)i / (float)(TABLE_ETA - 1); }

__device__ __forceinline__ float table_row(const float *t, uint32_t base, float r, float mu) {
    float x = clampf((mu - TABLE_MU_MIN) / (1.0f - TABLE_MU_MIN) * (float)(TABLE_MU - 1), 0.0f, (float)(TABLE_MU - 1));
    float y = clampf(r * (float)(TABLE_ROUGHNESS - 1), 0.0f, (float)(TABLE_ROUGHNESS - 1));
    uint32_t x0 = (uint32_t)x, y0 = (uint32_t)y;
    uint32_t x1 = min(x0 + 1u, TABLE_MU - 1), y1 = min(y0 + 1u, TABLE_ROUGHNESS - 1);
    float fx = x - (float)x0, fy = y - (float)y0;
    float a = mixf(t[base + y0 * TABLE_MU + x0], t[base + y0 * TABLE_MU + x1], fx);
    float b = mixf(t[base + y1 * TABLE_MU + x0], t[base + y1 * TABLE_MU + x1], fx);
    return mixf(a, b, fy);
}
__device__ __forceinline__ float table_ggx(const float *t, float mu, float r) { return table_row(t, TABLE_GGX, r, mu); }
__device__ __forceinline__ float table_dielectric(const float *t, float mu, float r, float eta) {
    float z = clampf(z_of_eta(eta) / TABLE_Z_MAX * (float)(TABLE_ETA - 1), 0.0f, (float)(TABLE_ETA - 1));
    uint32_t z0 = (uint32_t)z, z1 = min((uint32_t)z + 1u, TABLE_ETA - 1);
    uint32_t plane = TABLE_MU * TABLE_ROUGHNESS;
    return mixf(table_row(t, TABLE_DIELECTRIC + z0 * plane, r, mu), table_row(t, TABLE_DIELECTRIC + z1 * plane, r, mu),
                z - (float)z0);
}

// ------------------------------------------------------------------- EON

constexpr float FON_C1 = 0.5f - 2.0f / (3.0f * kPi);
constexpr float FON_C2 = 2.0f / 3.0f - 28.0f / (15.0f * kPi);

__device__ __forceinline__ float fon_albedo(float mu, float r) {
    float m = 1.0f - mu;
    float g_over_pi = m * (0.0571085289f + m * (0.491881867f + m * (-0.332181442f + m * 0.0714429953f)));
    return (1.0f + r * g_over_pi) / (1.0f + FON_C1 * r);
}

// Energy-preserving Oren–Nayar (Portsmouth, Kutz and Hill 2025).
__device__ __forceinline__ Spec eon(Spec rho, float r, V3 wi, V3 wo) {
    float mu_i = wi.z, mu_o = wo.z;
    float s = dot(wi, wo) - mu_i * mu_o;
    float s_over_t = s > 0.0f ? s / fmaxf(mu_i, mu_o) : s;
    float a = 1.0f / (1.0f + FON_C1 * r);
    Spec single = rho * (a * (1.0f + r * s_over_t) / kPi);
    float e_o = fon_albedo(mu_o, r), e_i = fon_albedo(mu_i, r);
    float average = a * (1.0f + FON_C2 * r);
    Spec rho_ms = rho * rho * average / (spec(1.0f) - rho * (1.0f - average));
    Spec multiple = rho_ms * (fmaxf(1e-7f, 1.0f - e_o) * fmaxf(1e-7f, 1.0f - e_i) / (kPi * fmaxf(1e-7f, 1.0f - average)));
    return single + multiple;
}

// F82-tint metal Fresnel (Kutz et al.).
__device__ __forceinline__ Spec fresnel_f82(Spec f0, Spec tint, float mu) {
    const float bar = 1.0f / 7.0f;
    float m = 1.0f - mu;
    float m5 = m * m * m * m * m;
    float b5 = powf(1.0f - bar, 5.0f);
    Spec schlick = f0 + (spec(1.0f) - f0) * m5;
    Spec schlick_bar = f0 + (spec(1.0f) - f0) * b5;
    Spec a = schlick_bar * (spec(1.0f) - tint) / (bar * b5 * (1.0f - bar));
    return sclamp(schlick - a * (mu * m5 * m), 0.0f, 1.0f);
}

__device__ __forceinline__ float rgb_luminance(V3 c) {
    return c.x * 0.2722287168f + c.y * 0.6740817658f + c.z * 0.0536895174f;
}

constexpr uint32_t LOBE_DIFFUSE = 0, LOBE_GLOSSY = 1, LOBE_TRANSMISSION = 2;

__device__ __forceinline__ float dispersed_ior(float ior, float b, float lambda) {
    return ior + b * (1.0f / (lambda * lambda) - 1.0f / (587.6f * 587.6f));
}

// ------------------------------------------------------------- the surface

struct Surface {
    V3 normal, tangent, bitangent, geometric;
    uint32_t material;
    bool back, inside;
    float diffuse_roughness, metalness, specular_weight, transmission, roughness, coat_roughness, ior, darkening;
    float lum_base, lum_specular, lum_transmission, lum_coat;
    float2 alpha;
    float eta, coat;
    float2 coat_alpha;
    float coat_eta;
    float dielectric_compensation, metal_missing, specular_albedo, coat_compensation, coat_albedo, coat_k;
    float p_coat, p_metal, p_specular, p_transmission, p_diffuse;
    Spec base, metal_tint, specular_tint, transmission_tint, coat_tint, substrate;
};

__device__ __forceinline__ V3 to_local(const Surface &s, V3 v) {
    return V3{dot(v, s.tangent), dot(v, s.bitangent), dot(v, s.normal)};
}
__device__ __forceinline__ V3 to_world(const Surface &s, V3 v) {
    return s.tangent * v.x + s.bitangent * v.y + s.normal * v.z;
}

__device__ __forceinline__ Surface surface_at(V3 normal, V3 geometric, uint32_t material, bool back) {
    Surface s;
    s.normal = normal;
    basis(normal, s.tangent, s.bitangent);
    s.geometric = geometric;
    s.material = material;
    s.back = back;
    return s;
}

// The OpenPBR parameters a material's node graph drives at one hit
// (openpbr.glsl:52-60): `mask` bit b set means the register group of
// shader_graph.lucb's surface table, values[] holding them. `viewed`: the
// program read the view direction (Fresnel or Layer Weight).
struct Program {
    uint32_t mask;
    bool viewed;
    float values[23];
    __device__ __forceinline__ bool driven(uint32_t bit) const { return (mask >> bit) & 1u; }
    __device__ __forceinline__ V3 driven3(uint32_t at) const { return V3{values[at], values[at + 1], values[at + 2]}; }
};

// openpbr.glsl:264-319: the material's parameters for the side seen; what
// a program drives (g, may be null) replaces the record's.
template <bool SPECTRAL>
__device__ __forceinline__ void read_material(const Params &p, Surface &s, float hero_wavelength,
                                              const Program *g = nullptr) {
    uint32_t m = s.material;
    float4 base_rgb = material_at(p, m, 1), specular_rgb = material_at(p, m, 3), coat_rgb = material_at(p, m, 5),
           transmission_rgb = material_at(p, m, 9), scalars = material_at(p, m, 10), indices = material_at(p, m, 11),
           geometry = material_at(p, m, 12);
    if (g && g->mask) {
        if (g->driven(0)) base_rgb.x = g->values[0], base_rgb.y = g->values[1], base_rgb.z = g->values[2];
        if (g->driven(1)) base_rgb.w = g->values[3];
        if (g->driven(2)) scalars.x = g->values[4];
        if (g->driven(3)) scalars.y = g->values[5];
        if (g->driven(4)) specular_rgb.w = g->values[6];
        if (g->driven(5)) scalars.z = g->values[7];
        if (g->driven(6)) coat_rgb.w = g->values[8];
        if (g->driven(7)) indices.y = g->values[9];
        if (g->driven(8)) transmission_rgb.w = g->values[10];
        if (g->driven(10)) specular_rgb.x = g->values[14], specular_rgb.y = g->values[15], specular_rgb.z = g->values[16];
        if (g->driven(11))
            transmission_rgb.x = g->values[17], transmission_rgb.y = g->values[18], transmission_rgb.z = g->values[19];
        if (g->driven(12)) coat_rgb.x = g->values[20], coat_rgb.y = g->values[21], coat_rgb.z = g->values[22];
    }
    s.diffuse_roughness = scalars.y;
    s.metalness = clampf(scalars.x, 0.0f, 1.0f);
    s.specular_weight = fmaxf(specular_rgb.w, 0.0f);
    s.transmission = clampf(transmission_rgb.w, 0.0f, 1.0f);
    s.inside = s.back && s.transmission > 0.0f;
    s.coat = !s.inside ? clampf(coat_rgb.w, 0.0f, 1.0f) : 0.0f;
    s.coat_roughness = clampf(indices.y, 0.0f, 1.0f);
    s.coat_eta = fmaxf(indices.z, 1.0001f);
    s.coat_alpha = make_float2(1.0f, 1.0f);
    if (s.coat > 0.0f) s.coat_alpha = ggx_alpha(s.coat_roughness, clampf(geometry.z, 0.0f, 1.0f));
    s.darkening = clampf(indices.w, 0.0f, 1.0f);
    float roughness = clampf(scalars.z, 0.0f, 1.0f);
    float r4 = roughness * roughness * roughness * roughness;
    float c4 = s.coat_roughness * s.coat_roughness * s.coat_roughness * s.coat_roughness;
    s.roughness = mixf(roughness, fminf(1.0f, powf(r4 + 2.0f * c4, 0.25f)), s.coat);
    s.alpha = ggx_alpha(s.roughness, clampf(scalars.w, 0.0f, 1.0f));
    float eta = fmaxf(indices.x, 1.0001f);
    if (SPECTRAL && geometry.w != 0.0f && hero_wavelength > 0.0f)
        eta = fmaxf(dispersed_ior(eta, geometry.w, hero_wavelength), 1.0001f);
    float relative = eta / s.coat_eta;
    if (relative < 1.0f) relative = 1.0f / relative;
    s.ior = fmaxf(mixf(eta, relative, s.coat), 1.0001f);
    s.eta = s.inside ? 1.0f / s.ior : s.ior;
    s.lum_base = fmaxf(rgb_luminance(V3{base_rgb.x * base_rgb.w, base_rgb.y * base_rgb.w, base_rgb.z * base_rgb.w}), 0.0f);
    s.lum_specular = fmaxf(rgb_luminance(v3(specular_rgb)), 0.0f);
    s.lum_transmission = fmaxf(rgb_luminance(v3(transmission_rgb)), 0.05f);
    s.lum_coat = rgb_luminance(v3(coat_rgb));
}

__device__ __forceinline__ void choose_lobes(Surface &s, float w_coat, float w_metal, float w_specular,
                                             float w_transmission, float w_diffuse) {
    if (s.inside) {
        w_metal = 0.0f;
        w_diffuse = 0.0f;
        w_specular = fmaxf(w_specular, 0.05f);
        w_transmission = fmaxf(w_transmission, s.transmission > 0.0f ? 0.05f : 0.0f);
    }
    float total = w_coat + w_metal + w_specular + w_transmission + w_diffuse;
    float scale = total > 0.0f ? 1.0f / total : 0.0f;
    s.p_coat = w_coat * scale;
    s.p_metal = w_metal * scale;
    s.p_specular = w_specular * scale;
    s.p_transmission = w_transmission * scale;
    s.p_diffuse = w_diffuse * scale;
}

// openpbr.glsl:324-370: albedos and lobe choice for the view direction wo.
__device__ __forceinline__ void orient_lobes(const Params &p, Surface &s, V3 wo) {
    float mu = clampf(wo.z, 0.0f, 1.0f);
    s.coat_compensation = 1.0f;
    s.coat_albedo = 0.0f;
    if (s.coat > 0.0f) {
        float e_coat = fmaxf(table_ggx(p.tables, mu, s.coat_roughness), 1e-4f);
        s.coat_compensation = 1.0f + fresnel_dielectric_average(s.coat_eta) * (1.0f - e_coat) / e_coat;
        s.coat_albedo = clampf(s.coat_compensation * table_dielectric(p.tables, mu, s.coat_roughness, s.coat_eta), 0.0f, 1.0f);
    }
    float e_ggx = fmaxf(table_ggx(p.tables, mu, s.roughness), 1e-4f);
    float missing = (1.0f - e_ggx) / e_ggx;
    s.metal_missing = missing;
    s.dielectric_compensation = 1.0f + fresnel_dielectric_average(s.ior) * missing;
    s.specular_albedo =
        clampf(s.specular_weight * s.dielectric_compensation * table_dielectric(p.tables, mu, s.roughness, s.ior), 0.0f, 1.0f);
    float substrate_rgb = 1.0f;
    s.coat_k = 0.0f;
    if (s.coat > 0.0f) {
        float k_smooth = fresnel_dielectric(mu, s.coat_eta);
        float k_rough = 1.0f - (1.0f - fresnel_dielectric_average(s.coat_eta)) / (s.coat_eta * s.coat_eta);
        float base_roughness = mixf(1.0f, s.roughness, s.metalness);
        s.coat_k = mixf(k_smooth, k_rough, base_roughness);
        float delta_rgb = (1.0f - s.coat_k) / fmaxf(1e-4f, 1.0f - s.lum_base * s.coat_k);
        substrate_rgb = mixf(1.0f, s.lum_coat * (1.0f - s.coat_albedo) * mixf(1.0f, delta_rgb, s.darkening), s.coat);
    }
    float dielectric = 1.0f - s.metalness;
    choose_lobes(s, s.coat * s.coat_albedo, substrate_rgb * s.metalness,
                 substrate_rgb * dielectric * s.specular_albedo * s.lum_specular,
                 substrate_rgb * dielectric * s.transmission * (1.0f - s.specular_albedo) * s.lum_transmission,
                 substrate_rgb * dielectric * (1.0f - s.transmission) * (1.0f - s.specular_albedo) * s.lum_base);
}

__device__ __forceinline__ void update_substrate(Surface &s) {
    s.substrate = spec(1.0f);
    if (s.coat > 0.0f) {
        Spec delta = spec(1.0f - s.coat_k) / (spec(1.0f) - s.base * s.coat_k);
        Spec under = s.coat_tint * (1.0f - s.coat_albedo) * mixs(spec(1.0f), delta, s.darkening);
        s.substrate = mixs(spec(1.0f), under, s.coat);
    }
}

// A color a program drives: a reflectance through the environment basis, at
// most one (openpbr.glsl:391-401).
template <bool SPECTRAL> __device__ __forceinline__ Spec driven_color(const Params &p, V3 rgb, Spec lambda) {
    Spec c = spec_of_texel<SPECTRAL>(p, rgb, lambda);
    return make_float4(clampf(c.x, 0.0f, 1.0f), clampf(c.y, 0.0f, 1.0f), clampf(c.z, 0.0f, 1.0f), clampf(c.w, 0.0f, 1.0f));
}

template <bool SPECTRAL>
__device__ __forceinline__ void read_colors(const Params &p, Surface &s, Spec lambda, const Program *g = nullptr) {
    uint32_t m = s.material;
    if (g && g->mask) {
        float weight = g->driven(1) ? g->values[3] : material_at(p, m, 1).w;
        s.base = (g->driven(0) ? driven_color<SPECTRAL>(p, g->driven3(0), lambda) : material_color<SPECTRAL>(p, m, 0, lambda)) *
                 weight;
        s.metal_tint =
            g->driven(10) ? driven_color<SPECTRAL>(p, g->driven3(14), lambda) : material_color<SPECTRAL>(p, m, 2, lambda);
        s.specular_tint = s.metal_tint * s.specular_weight;
        s.transmission_tint =
            g->driven(11) ? driven_color<SPECTRAL>(p, g->driven3(17), lambda) : material_color<SPECTRAL>(p, m, 8, lambda);
        s.coat_tint = s.coat > 0.0f ? (g->driven(12) ? driven_color<SPECTRAL>(p, g->driven3(20), lambda)
                                                     : material_color<SPECTRAL>(p, m, 4, lambda))
                                    : spec(1.0f);
        return;
    }
    float weight = material_at(p, m, 1).w;
    s.base = material_color<SPECTRAL>(p, m, 0, lambda) * weight;
    s.metal_tint = material_color<SPECTRAL>(p, m, 2, lambda);
    s.specular_tint = s.metal_tint * s.specular_weight;
    s.transmission_tint = material_color<SPECTRAL>(p, m, 8, lambda);
    s.coat_tint = s.coat > 0.0f ? material_color<SPECTRAL>(p, m, 4, lambda) : spec(1.0f);
}

template <bool SPECTRAL>
__device__ __forceinline__ void prepare_surface(const Params &p, Surface &s, V3 wo, Spec lambda, float hero,
                                                const Program *g = nullptr) {
    read_material<SPECTRAL>(p, s, hero, g);
    orient_lobes(p, s, wo);
    read_colors<SPECTRAL>(p, s, lambda, g);
    update_substrate(s);
}

__device__ __forceinline__ void orient_surface(const Params &p, Surface &s, V3 wo) {
    orient_lobes(p, s, wo);
    update_substrate(s);
}

// The lobe choice light paths sample with: Fresnel at normal incidence for
// the albedos, so it holds for every direction (openpbr.glsl:392-420).
__device__ __forceinline__ void light_lobes(Surface &s) {
    float f0 = (s.ior - 1.0f) / (s.ior + 1.0f);
    float specular_albedo = clampf(s.specular_weight * f0 * f0, 0.0f, 1.0f);
    float coat_albedo = 0.0f;
    if (s.coat > 0.0f) {
        float c0 = (s.coat_eta - 1.0f) / (s.coat_eta + 1.0f);
        coat_albedo = c0 * c0;
    }
    float substrate_rgb = mixf(1.0f, s.lum_coat * (1.0f - coat_albedo), s.coat);
    float dielectric = 1.0f - s.metalness;
    choose_lobes(s, s.coat * coat_albedo, substrate_rgb * s.metalness,
                 substrate_rgb * dielectric * specular_albedo * s.lum_specular,
                 substrate_rgb * dielectric * s.transmission * (1.0f - specular_albedo) * s.lum_transmission,
                 substrate_rgb * dielectric * (1.0f - s.transmission) * (1.0f - specular_albedo) * s.lum_base);
}

__device__ __forceinline__ bool refraction_half(V3 wi, V3 wo, float eta, V3 &h) {
    h = normalize(wo + wi * eta);
    if (h.z < 0.0f) h = -h;
    return dot(wo, h) > 0.0f && dot(wi, h) < 0.0f;
}

__device__ __forceinline__ float surface_pdf(const Surface &s, V3 wi, V3 wo) {
    if (wo.z <= 0.0f || wi.z == 0.0f) return 0.0f;
    if (wi.z < 0.0f) {
        if (s.p_transmission <= 0.0f) return 0.0f;
        V3 h;
        if (!refraction_half(wi, wo, s.eta, h)) return 0.0f;
        float o_h = dot(wo, h), i_h = dot(wi, h);
        float den = o_h + s.eta * i_h;
        float jacobian = s.eta * s.eta * fabsf(i_h) / (den * den);
        return s.p_transmission * ggx_g1(wo, s.alpha) * ggx_d(h, s.alpha) * o_h / wo.z * jacobian;
    }
    V3 h = normalize(wi + wo);
    float pdf = 0.0f;
    if (s.p_coat > 0.0f) pdf += s.p_coat * ggx_reflection_pdf(wo, h, s.coat_alpha);
    if (!s.inside) pdf += s.p_diffuse * wi.z / kPi;
    pdf += (s.p_specular + s.p_metal) * ggx_reflection_pdf(wo, h, s.alpha);
    return pdf;
}

// openpbr.glsl:481-517: f (cosine not included) and the pdf of wi.
__device__ __forceinline__ Spec evaluate_surface(const Surface &s, V3 wi, V3 wo, float &pdf) {
    pdf = 0.0f;
    if (wo.z <= 0.0f || wi.z == 0.0f) return spec(0.0f);
    pdf = surface_pdf(s, wi, wo);
    float dielectric = 1.0f - s.metalness;
    if (wi.z < 0.0f) {
        if (s.p_transmission <= 0.0f) return spec(0.0f);
        V3 h;
        if (!refraction_half(wi, wo, s.eta, h)) return spec(0.0f);
        float o_h = dot(wo, h), i_h = dot(wi, h);
        float den = o_h + s.eta * i_h;
        float d = ggx_d(h, s.alpha);
        float f = (1.0f - fresnel_dielectric(o_h, s.eta)) * d * ggx_g2(-wi, wo, s.alpha) * fabsf(i_h) * o_h /
                  (fabsf(wi.z) * wo.z * den * den);
        return s.substrate * s.transmission_tint * (dielectric * s.transmission * f);
    }
    V3 h = normalize(wi + wo);
    float cos_h = dot(wo, h);
    Spec f = spec(0.0f);
    if (s.p_coat > 0.0f) {
        float coat = ggx_d(h, s.coat_alpha) * ggx_g2(wi, wo, s.coat_alpha) / (4.0f * wi.z * wo.z);
        f = f + spec(s.coat * fresnel_dielectric(cos_h, s.coat_eta) * s.coat_compensation * coat);
    }
    Spec below = spec(0.0f);
    if (!s.inside)
        below = below + eon(s.base, s.diffuse_roughness, wi, wo) *
                            (dielectric * (1.0f - s.transmission) * (1.0f - s.specular_albedo));
    float microfacet = ggx_d(h, s.alpha) * ggx_g2(wi, wo, s.alpha) / (4.0f * wi.z * wo.z);
    if (dielectric > 0.0f)
        below = below + s.specular_tint * (dielectric * fresnel_dielectric(cos_h, s.eta) * s.dielectric_compensation * microfacet);
    if (s.metalness > 0.0f && !s.inside) {
        Spec average = (s.base * 20.0f + spec(1.0f)) / 21.0f;
        Spec compensation = spec(1.0f) + average * s.metal_missing;
        below = below + smin(spec(1.0f), fresnel_f82(s.base, s.metal_tint, cos_h) * s.specular_weight) * compensation *
                            (s.metalness * microfacet);
    }
    return f + s.substrate * below;
}

// openpbr.glsl:522-547: the lobe and the direction, without the density.
__device__ __forceinline__ bool pick_direction(const Surface &s, V3 wo, float ux, float uy, float uz, V3 &wi,
                                               uint32_t &lobe) {
    if (wo.z <= 0.0f) return false;
    float pick = uz;
    if (pick < s.p_diffuse) {
        float2 d = concentric(ux, uy);
        wi = V3{d.x, d.y, sqrtf(fmaxf(0.0f, 1.0f - (d.x * d.x + d.y * d.y)))};
        lobe = LOBE_DIFFUSE;
    } else if ((pick -= s.p_diffuse) < s.p_coat) {
        wi = reflect(-wo, ggx_sample(wo, s.coat_alpha, ux, uy));
        lobe = LOBE_GLOSSY;
    } else if ((pick -= s.p_coat) < s.p_specular + s.p_metal) {
        wi = reflect(-wo, ggx_sample(wo, s.alpha, ux, uy));
        lobe = LOBE_GLOSSY;
    } else if (s.p_transmission > 0.0f) {
        V3 h = ggx_sample(wo, s.alpha, ux, uy);
        wi = refract(-wo, h, 1.0f / s.eta);
        if (dot(wi, wi) == 0.0f) return false;
        lobe = LOBE_TRANSMISSION;
    } else {
        return false;
    }
    return wi.z != 0.0f;
}

// openpbr.glsl:522-547
__device__ __forceinline__ bool sample_direction(const Surface &s, V3 wo, float ux, float uy, float uz, V3 &wi,
                                                 float &pdf, uint32_t &lobe) {
    pdf = 0.0f;
    if (!pick_direction(s, wo, ux, uy, uz, wi, lobe)) return false;
    pdf = surface_pdf(s, wi, wo);
    return pdf > 0.0f;
}

__device__ __forceinline__ bool sample_surface(const Surface &s, V3 wo, float ux, float uy, float uz, V3 &wi,
                                               Spec &weight, float &pdf, uint32_t &lobe) {
    if (!sample_direction(s, wo, ux, uy, uz, wi, pdf, lobe)) return false;
    float again;
    Spec f = evaluate_surface(s, wi, wo, again);
    weight = f * (fabsf(wi.z) / pdf);
    return true;
}

}  // namespace rt
}  // namespace lr
