// Materials' node-graph programs on the device (shaders/shader_program.glsl;
// the compiler is luce_render/shader_graph.py after shader_graph.lucb).
//
// A material whose record's row 13 x is >= 0 runs the program at that word
// once a hit, before its parameters are read: a flat list of u32 words, an
// opcode then operands and output stack slots. An operand is a constant's f32
// bits or a stack slot tagged in a NaN (0x7FC10000 | slot). The interpreter
// keeps its 64-value stack in private memory; word reads stop at the words'
// end and stack indices stay inside the stack, so a malformed program ends
// early or computes garbage but never reads or writes outside its buffers.
#pragma once

#include "render_bsdf.h"

namespace lr {
namespace rt {

constexpr uint32_t kProgramStack = 64u;
constexpr uint32_t kSlotTag = 0x7FC10000u;
constexpr uint32_t kProgramSteps = 4096u;
constexpr uint32_t kDrivenValues = 23u;

// The texture coordinates of a hit at barycentrics (b1, b2) (openpbr.glsl:41-50).
__device__ __forceinline__ float2 triangle_uv(const Params &p, uint32_t tri, float b1, float b2) {
    const float *t = p.uvs + (size_t)tri * 6u;
    float w = 1.0f - b1 - b2;
    return make_float2(t[0] * w + t[2] * b1 + t[4] * b2, t[1] * w + t[3] * b1 + t[5] * b2);
}

// A texture slot's bilinear sample at uv, repeating, v = 0 the first row
// (a linear, repeat-addressed sampler over fract(uv), openpbr.glsl:66-68).
__device__ __forceinline__ float4 texture_at(const Params &p, uint32_t slot, float u, float v) {
    if (slot >= p.texture_count) return make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    uint4 t = p.texture_table[slot];
    if (t.y == 0u || t.z == 0u) return make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    int w = (int)t.y, h = (int)t.z;
    float x = fractf(u) * (float)w - 0.5f, y = fractf(v) * (float)h - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    // Non-finite coordinates sample the first texel.
    if (!(fx >= -1.0f && fx <= (float)w)) fx = x = 0.0f;
    if (!(fy >= -1.0f && fy <= (float)h)) fy = y = 0.0f;
    float ax = x - fx, ay = y - fy;
    int x0 = (int)fx, y0 = (int)fy;
    int x1 = x0 + 1, y1 = y0 + 1;
    x0 = x0 < 0 ? x0 + w : (x0 >= w ? x0 - w : x0);
    x1 = x1 < 0 ? x1 + w : (x1 >= w ? x1 - w : x1);
    y0 = y0 < 0 ? y0 + h : (y0 >= h ? y0 - h : y0);
    y1 = y1 < 0 ? y1 + h : (y1 >= h ? y1 - h : y1);
    const float4 *texels = p.textures + t.x;
    float4 a = texels[(size_t)y0 * w + x0], b = texels[(size_t)y0 * w + x1];
    float4 c = texels[(size_t)y1 * w + x0], d = texels[(size_t)y1 * w + x1];
    auto lerp4 = [](float4 l, float4 r, float s) {
        return make_float4(mixf(l.x, r.x, s), mixf(l.y, r.y, s), mixf(l.z, r.z, s), mixf(l.w, r.w, s));
    };
    return lerp4(lerp4(a, b, ax), lerp4(c, d, ax), ay);
}

// One of the cube's twelve edge directions, by a hash of the lattice point
// (Perlin's improved noise; shader_program.glsl:27-35).
__device__ __forceinline__ V3 gradient_at(int cx, int cy, int cz) {
    uint32_t h = hash_u32((uint32_t)cx * 73856093u ^
                          hash_u32((uint32_t)cy * 19349663u ^ hash_u32((uint32_t)cz * 83492791u))) % 12u;
    float sx = (h & 1u) ? -1.0f : 1.0f, sy = (h & 2u) ? -1.0f : 1.0f;
    if (h < 4u) return V3{sx, sy, 0.0f};
    if (h < 8u) return V3{sx, 0.0f, sy};
    return V3{0.0f, sx, sy};
}

// Perlin gradient noise in about -1..1.
__device__ __noinline__ float gradient_noise(V3 q) {
    V3 i{floorf(q.x), floorf(q.y), floorf(q.z)};
    V3 f = q - i;
    auto fade = [](float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); };
    float ux = fade(f.x), uy = fade(f.y), uz = fade(f.z);
    int cx = (int)i.x, cy = (int)i.y, cz = (int)i.z;
    float n[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
        n[k] = dot(gradient_at(cx + dx, cy + dy, cz + dz), f - V3{(float)dx, (float)dy, (float)dz});
    }
    return mixf(mixf(mixf(n[0], n[1], ux), mixf(n[2], n[3], ux), uy), mixf(mixf(n[4], n[5], ux), mixf(n[6], n[7], ux), uy),
                uz);
}

// Fractal noise in 0..1: `detail` octaves (a fractional one blends in), each
// `roughness` as strong as the last.
__device__ __forceinline__ float fractal_noise(V3 q, float detail, float roughness) {
    float octaves = clampf(detail, 0.0f, 15.0f);
    int whole = (int)floorf(octaves);
    float part = octaves - (float)whole;
    float sum = 0.0f, total = 0.0f, amplitude = 1.0f;
    for (int octave = 0; octave <= whole; octave++) {
        sum += gradient_noise(q) * amplitude;
        total += amplitude;
        amplitude *= clampf(roughness, 0.0f, 1.0f);
        q = q * 2.0f;
    }
    if (part > 0.0f) {
        sum += gradient_noise(q) * amplitude * part;
        total += amplitude * part;
    }
    return clampf(0.5f + 0.5f * sum / total, 0.0f, 1.0f);
}

__device__ __forceinline__ float math_of(uint32_t operation, float a, float b) {
    switch (operation) {
        case 0u: return a + b;
        case 1u: return a - b;
        case 2u: return a * b;
        case 3u: return b != 0.0f ? a / b : 0.0f;
        case 4u:
            if (a >= 0.0f) return powf(a, b);
            if (b != floorf(b)) return 0.0f;
            return powf(-a, b) * (b - 2.0f * floorf(b * 0.5f) == 1.0f ? -1.0f : 1.0f);
        case 5u: return fminf(a, b);
        case 6u: return fmaxf(a, b);
        case 7u: return fabsf(a);
        case 8u: return sinf(a);
        case 9u: return cosf(a);
        case 10u: return floorf(a);
        default: return fractf(a);
    }
}

// Two unit tangents about n (Duff et al.'s branchless frame).
__device__ __forceinline__ void bump_tangents(V3 n, V3 &t, V3 &b) {
    float s = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (s + n.z);
    float c = n.x * n.y * a;
    t = V3{1.0f + s * n.x * n.x * a, s * c, -s * n.x};
    b = V3{c, s + n.y * n.y * a, -n.y};
}

// The bump footprint: the camera pixel's width at x, `pixels` of them, with
// a floor above f32's resolution of x (shader_program.glsl:121-130).
__device__ __forceinline__ float bump_width(const Params &p, V3 x, float pixels) {
    V3 du{p.frame[8], p.frame[9], p.frame[10]};
    float pixel = length(du) * (p.frame[3] > 0.5f ? 1.0f : distance(x, V3{p.frame[0], p.frame[1], p.frame[2]}));
    float m = fmaxf(fabsf(x.x), fmaxf(fabsf(x.y), fabsf(x.z)));
    return fmaxf(pixel * fmaxf(pixels, 0.0f), 4e-6f * (1.0f + m));
}

// How the texture coordinates change along a step in the triangle's plane
// (zero for a degenerate triangle): the step's barycentrics by the dual basis.
__device__ __forceinline__ float2 uv_step(const Params &p, uint32_t tri, V3 step) {
    V3 p0, e1, e2;
    triangle_edges(p, tri, p0, e1, e2);
    float g11 = dot(e1, e1), g12 = dot(e1, e2), g22 = dot(e2, e2);
    float det = g11 * g22 - g12 * g12;
    if (fabsf(det) < 1e-24f) return make_float2(0.0f, 0.0f);
    float a = (g22 * dot(e1, step) - g12 * dot(e2, step)) / det;
    float b = (g11 * dot(e2, step) - g12 * dot(e1, step)) / det;
    const float *t = p.uvs + (size_t)tri * 6u;
    return make_float2((t[2] - t[0]) * a + (t[4] - t[0]) * b, (t[3] - t[1]) * a + (t[5] - t[1]) * b);
}

// A tangent-space normal map's color on the hit's triangle, about the
// shading normal n in the frame its texture coordinates give.
__device__ __forceinline__ V3 normal_from_map(const Params &p, uint32_t tri, V3 n, V3 color, float strength) {
    V3 p0, e1, e2;
    triangle_edges(p, tri, p0, e1, e2);
    const float *q = p.uvs + (size_t)tri * 6u;
    float d1x = q[2] - q[0], d1y = q[3] - q[1], d2x = q[4] - q[0], d2y = q[5] - q[1];
    float det = d1x * d2y - d2x * d1y;
    if (fabsf(det) < 1e-12f) return n;
    V3 t = (e1 * d2y - e2 * d1y) / det;
    V3 b = (e2 * d1x - e1 * d2x) / det;
    t = t - n * dot(n, t);
    if (dot(t, t) < 1e-20f) return n;
    t = normalize(t);
    V3 bt = cross(n, t);
    if (dot(bt, b) < 0.0f) bt = -bt;
    V3 c = color * 2.0f - V3{1.0f, 1.0f, 1.0f};
    V3 bent = t * c.x + bt * c.y + n * c.z;
    if (dot(bent, bent) < 1e-12f) return n;
    V3 u = normalize(bent);
    float s = clampf(strength, 0.0f, 1.0f);
    return normalize(n * (1.0f - s) + u * s);
}

// Cycles' bump (svm_node_set_bump): the surface gradient of the heights about
// `normal` (zero: n), `distance` long, blended in by `strength`.
__device__ __forceinline__ V3 bumped(const Params &p, V3 x, V3 n, V3 normal, float strength, float distance_scale,
                                     float pixels, float hc, float hx, float hy) {
    V3 base = dot(normal, normal) > 1e-12f ? normalize(normal) : n;
    V3 t, b;
    bump_tangents(n, t, b);
    float w = bump_width(p, x, pixels);
    V3 dx = t * w, dy = b * w;
    V3 rx = cross(dy, base), ry = cross(base, dx);
    float det = dot(dx, rx);
    V3 gradient = rx * (hx - hc) + ry * (hy - hc);
    float sign = det > 0.0f ? 1.0f : (det < 0.0f ? -1.0f : 0.0f);
    V3 out = base * fabsf(det) - gradient * (distance_scale * sign);
    if (dot(out, out) < 1e-30f) return base;
    V3 u = normalize(out);
    float s = clampf(strength, 0.0f, 1.0f);
    return normalize(base * (1.0f - s) + u * s);
}

// Facing: Cycles' Fresnel and Layer Weight about `normal` (zero: n), a view
// on the far side meaning the surface is seen from behind.
__device__ __forceinline__ float facing_cos(V3 n, V3 normal, V3 view) {
    return dot(view, dot(normal, normal) > 1e-12f ? normalize(normal) : n);
}

__device__ __forceinline__ float fresnel_node(V3 n, V3 normal, V3 view, float ior) {
    float c = facing_cos(n, normal, view);
    float eta = fmaxf(ior, 1e-5f);
    return fresnel_dielectric(c, c < 0.0f ? 1.0f / eta : eta);
}

__device__ __forceinline__ float2 layer_weight_node(V3 n, V3 normal, V3 view, float blend) {
    float c = facing_cos(n, normal, view);
    float eta = fmaxf(1.0f - blend, 1e-5f);
    float fresnel = fresnel_dielectric(c, c < 0.0f ? eta : 1.0f / eta);
    float facing = fabsf(c);
    if (blend != 0.5f) {
        float b = clampf(blend, 0.0f, 1.0f - 1e-5f);
        facing = powf(facing, b < 0.5f ? 2.0f * b : 0.5f / (1.0f - b));
    }
    return make_float2(fresnel, 1.0f - facing);
}

// Runs material m's program at the hit on `tri`: point x, outward shading
// normal n, `view` the unit direction toward the camera's side of the path,
// texture coordinates uv. Fills g (shader_program.glsl:205-349).
__device__ __noinline__ void run_program(const Params &p, Program &g, uint32_t m, uint32_t tri, V3 x, V3 n, V3 view,
                                         float2 uv) {
    g.mask = 0u;
    g.viewed = false;
    float entry = material_at(p, m, 13).x;
    if (!(entry >= 0.0f) || !(entry < (float)p.program_words)) return;
    float stack[kProgramStack];
    const uint32_t *code = p.programs;
    const uint32_t words = p.program_words;
    uint32_t pc = (uint32_t)entry;
    auto word = [&]() -> uint32_t { return pc < words ? code[pc++] : 0u; };
    auto operand = [&]() -> float {
        uint32_t w = word();
        return (w & 0xFFFF0000u) == kSlotTag ? stack[w & (kProgramStack - 1u)] : __uint_as_float(w);
    };
    auto operand3 = [&]() -> V3 {
        float a = operand(), b = operand(), c = operand();
        return V3{a, b, c};
    };
    auto put = [&](float v) { stack[word() & (kProgramStack - 1u)] = v; };
    auto put3 = [&](V3 v) {
        put(v.x);
        put(v.y);
        put(v.z);
    };
    for (uint32_t step = 0; step < kProgramSteps; step++) {
        if (pc >= words) return;
        uint32_t op = code[pc++];
        switch (op) {
            case 0u: return;
            case 1u:
            case 2u: {
                // A bump's offset copies read the point a step away.
                uint32_t axis = word();
                float pixels = operand();
                V3 offset{0.0f, 0.0f, 0.0f};
                if (axis != 0u) {
                    V3 t, b;
                    bump_tangents(n, t, b);
                    offset = (axis == 1u ? t : b) * bump_width(p, x, pixels);
                }
                V3 v;
                if (op == 1u) {
                    float2 d = axis != 0u ? uv_step(p, tri, offset) : make_float2(0.0f, 0.0f);
                    v = V3{uv.x + d.x, uv.y + d.y, 0.0f};
                } else {
                    v = x + offset;
                }
                put3(v);
                break;
            }
            case 3u: {
                float v = operand();
                put(v);
                break;
            }
            case 4u:
            case 12u:
            case 13u: put3(operand3()); break;
            case 5u: {
                uint32_t slot = word();
                bool color = word() != 0u;
                V3 at = operand3();
                float4 texel = texture_at(p, slot, at.x, at.y);
                V3 rgb{texel.x, texel.y, texel.z};
                if (color)
                    rgb = V3{0.613097f * texel.x + 0.339523f * texel.y + 0.047379f * texel.z,
                             0.070194f * texel.x + 0.916354f * texel.y + 0.013452f * texel.z,
                             0.020616f * texel.x + 0.109570f * texel.y + 0.869815f * texel.z};
                put3(rgb);
                put(texel.w);
                break;
            }
            case 6u: {
                V3 at = operand3(), a = operand3(), b = operand3();
                float scale = operand();
                // A small offset keeps faces at whole coordinates off the seams.
                int cx = (int)floorf(at.x * scale + 1e-5f), cy = (int)floorf(at.y * scale + 1e-5f),
                    cz = (int)floorf(at.z * scale + 1e-5f);
                float fac = (float)((cx + cy + cz) & 1);
                put3(fac > 0.5f ? a : b);
                put(fac);
                break;
            }
            case 7u: {
                V3 at = operand3();
                float scale = operand(), detail = operand(), roughness = operand();
                V3 q = at * scale;
                float fac = fractal_noise(q, detail, roughness);
                float g1 = fractal_noise(q + V3{19.1f, 33.4f, 47.2f}, detail, roughness);
                float g2 = fractal_noise(q + V3{73.7f, 11.3f, 29.9f}, detail, roughness);
                put(fac);
                put3(V3{fac, g1, g2});
                break;
            }
            case 8u: {
                uint32_t operation = word();
                float a = operand(), b = operand();
                put(math_of(operation, a, b));
                break;
            }
            case 9u: {
                float fac = clampf(operand(), 0.0f, 1.0f);
                V3 a = operand3(), b = operand3();
                put3(V3{mixf(a.x, b.x, fac), mixf(a.y, b.y, fac), mixf(a.z, b.z, fac)});
                break;
            }
            case 10u: {
                float v = operand(), from_min = operand(), from_max = operand(), to_min = operand(), to_max = operand();
                float t = from_max != from_min ? (v - from_min) / (from_max - from_min) : 0.0f;
                put(to_min + t * (to_max - to_min));
                break;
            }
            case 11u: {
                float v = operand(), low = operand(), high = operand();
                put(clampf(v, fminf(low, high), fmaxf(low, high)));
                break;
            }
            case 14u: {
                float fac = operand();
                V3 a = operand3(), b = operand3();
                float first = operand(), last = operand();
                float t = last != first ? clampf((fac - first) / (last - first), 0.0f, 1.0f) : (fac < first ? 0.0f : 1.0f);
                put3(V3{mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)});
                break;
            }
            case 15u: {
                V3 color = operand3();
                float strength = operand();
                put3(normal_from_map(p, tri, n, color, strength));
                break;
            }
            case 16u: {
                uint32_t first = word(), width = word(), bit = word();
                for (uint32_t k = 0; k < width && k < kDrivenValues; k++) {
                    float v = operand();
                    if (first + k < kDrivenValues) g.values[first + k] = v;
                }
                if (bit < 32u) g.mask |= 1u << bit;
                break;
            }
            case 17u: {
                float strength = operand(), distance_scale = operand(), pixels = operand();
                float hc = operand(), hx = operand(), hy = operand();
                V3 normal = operand3();
                put3(bumped(p, x, n, normal, strength, distance_scale, pixels, hc, hx, hy));
                break;
            }
            case 18u:
            case 19u: {
                float parameter = operand();
                V3 normal = operand3();
                g.viewed = true;
                if (op == 18u) {
                    put(fresnel_node(n, normal, view, parameter));
                } else {
                    float2 w = layer_weight_node(n, normal, view, parameter);
                    put(w.x);
                    put(w.y);
                }
                break;
            }
            default: return;
        }
    }
}

}  // namespace rt
}  // namespace lr
