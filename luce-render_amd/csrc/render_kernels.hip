// The wavefront integrator's kernels (shaders/camera.comp, light_emit.comp,
// intersect_closest.comp, shade_surface.comp with light_paths.glsl,
// shade_miss.comp, intersect_shadow.comp, finish.comp, guide_fit.comp,
// film_convert.comp, tables.comp).
//
// MI355X form: every kernel is a persistent grid-stride loop over its queue,
// whose length it reads from device memory, so a pass is a fixed sequence of
// launches (no host round trip, no indirect dispatch) that records once into a
// hipGraph and replays. Queue pushes are wave-aggregated (one atomic per wave).
// The queue a kernel consumed is emptied by a later kernel of the same bounce
// that neither reads nor fills it. Pass-to-pass values (sample, band) live in
// a device PassState that begin_pass / end_pass advance.
#include "render_kernels.h"

#include "render_bsdf.h"
#include "render_guide.h"
#include "render_program.h"
#include "render_splats.h"
#include "trace_kernels.h"

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

// ------------------------------------------------------------ pass control

// Per-bounce grids are sized for every slot; once few paths remain most of
// their workgroups have no item. With one request such a workgroup leaves
// after one load of the queue's count, before the row plan (workgroup 0
// stays: kernels clear counters there). It would have run no iteration.
template <class F> __device__ __forceinline__ bool idle_block(const Params *plist, uint32_t n, F count_of) {
    if (n != 1u || blockIdx.x == 0u) return false;
    return blockIdx.x * blockDim.x >= count_of(plist[0]);
}

__global__ void begin_pass_kernel(const Params *plist) {
    const Params &p = plist[blockIdx.x];
    PassState *s = p.pass;
    if (threadIdx.x < 16) p.queues[threadIdx.x] = 0u;
    if (threadIdx.x == 0) {
        uint32_t start = s->band * p.band_size;
        uint32_t remaining = p.pixels - start;
        s->band_start = start;
        s->band_count = remaining < p.band_size ? remaining : p.band_size;
        s->light_count = s->band == 0 ? p.light_paths : 0u;
        // Samples this pass: one a band when a sample takes several, else up
        // to pass_samples, stopping at the target (the host counts alike).
        uint32_t left = p.sample_target > s->sample ? p.sample_target - s->sample : 1u;
        uint32_t k = s->bands == 1u ? (p.pass_samples < left ? p.pass_samples : left) : 1u;
        if (p.learned && s->sample < p.train_samples) k = 1u;  // it trains after each of those samples
        s->sub_samples = k > 0u ? k : 1u;
        if (s->light_count > 0) s->light_samples += s->sub_samples;
    }
}

__global__ void end_pass_kernel(const Params *plist) {
    const Params &p = plist[blockIdx.x];
    PassState *s = p.pass;
    if (threadIdx.x == 0) {
        s->band += 1u;
        if (s->band == s->bands) {
            s->band = 0;
            s->sample += s->sub_samples;
            s->samples_done = s->sample;
        }
    }
}

// Items of a persistent loop: wave-uniform trip counts, so every lane reaches
// the wave-wide pushes together.
#define FOR_ITEMS(count, i, active)                                                                         \
    for (uint32_t _base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); _base < (count);                 \
         _base += gridDim.x * blockDim.x)                                                                   \
        for (uint32_t i = _base + (threadIdx.x & 63u), _once = 1; _once; _once = 0)                         \
            for (bool active = i < (count); _once; _once = 0)

// The same with workgroup-uniform trip counts, for loops whose body ends in a
// workgroup-wide StagedPushes::sync.
#define FOR_BLOCK_ITEMS(count, i, active)                                                                   \
    for (uint32_t _base = blockIdx.x * blockDim.x; _base < (count); _base += gridDim.x * blockDim.x)        \
        for (uint32_t i = _base + threadIdx.x, _once = 1; _once; _once = 0)                                 \
            for (bool active = i < (count); _once; _once = 0)

// Staged ids per queue: up to eight iterations of a workgroup between flushes.
constexpr uint32_t kStageIterations = 8u;

// --------------------------------------------------------------- camera

// The pixel of the i-th camera path of a sample: the image is walked in 8x8
// tiles (a row of tiles at a time; the last row and column of tiles may be
// smaller), so a wave's 64 rays leave one tile together and meet the same
// nodes. Each pixel still draws its own samples (sample4 keys on the pixel),
// so images do not depend on the walk; the reference walks rows
// (camera.comp:14-20).
__device__ __forceinline__ uint32_t tiled_pixel(uint32_t i, uint32_t width, uint32_t height) {
    uint32_t tile_row = i / (8u * width), local = i - tile_row * 8u * width;
    uint32_t rows = min(8u, height - 8u * tile_row);
    uint32_t tile = local / (8u * rows), within = local - tile * 8u * rows;
    uint32_t cols = min(8u, width - 8u * tile);
    return (8u * tile_row + within / cols) * width + 8u * tile + within % cols;
}

template <bool SPECTRAL> __global__ __launch_bounds__(256) void camera_kernel(const Params *plist, uint32_t n) {
    __shared__ RowPlan plan;
    uint32_t total = plan.make(plist, n, [](const Params &q) { return pass_cameras(*q.pass); });
    __shared__ StagedPushes<1, 256 * kStageIterations> staged;
    staged.init({Q_CLOSEST});
    FOR_ROW_ITEMS(plan, n, total, g, r, path, active) {
        const Params &p = plist[r];
        State st(p);
        const PassState &ps = *p.pass;
        bool push = false;
        if (active) {
            SlotIdentity who = slot_identity(ps, path);
            uint32_t pixel = tiled_pixel(ps.band_start + who.slot, p.width, p.height);
            if (p.film[(size_t)FILM_HALF * p.pixels + pixel].w > 0.5f) {
                st.set_u(SU_PIXEL, path, kNoneU);
            } else {
                uint32_t width = p.width;
                float4 u = sample4(p.seed, pixel, who.sample, GROUP_CAMERA);
                float fx = (float)(pixel % width) + u.x, fy = (float)(pixel / width) + u.y;
                V3 film = frame_corner(p) + frame_du(p) * fx + frame_dv(p) * fy;
                bool ortho = p.frame[3] > 0.5f;
                V3 origin = ortho ? frame_position(p) + film : frame_position(p);
                V3 direction = normalize(ortho ? cross(frame_du(p), frame_dv(p)) : film);
                if (!ortho && K(p, K_CAMERA).z > 0.0f) {
                    V3 focused = frame_position(p) + direction * (K(p, K_CAMERA).w / dot(direction, v3(K(p, K_VIEW))));
                    origin = lens_point(p, u.z, u.w);
                    direction = normalize(focused - origin);
                }
                st.set_u(SU_PIXEL, path, pixel);
                st.set_u(SU_BOUNCE, path, 0u);
                st.set_v(SV_ORIGIN, path, f4(origin, 0.0f));
                st.set_v(SV_DIRECTION, path, f4(direction, 0.0f));
                st.set_v(SV_THROUGHPUT, path, spec(1.0f));
                st.set_v(SV_RADIANCE, path, spec(0.0f));
                st.set_v(SV_VERTEX, path, f4(origin, 1.0f));
                if (light_tracing(p))
                    st.set_v(SV_MIS, path, make_float4(light_paths_per_sample(p), 0.0f, 0.0f, camera_pdf(p, direction)));
                for (uint32_t k = 0; k < 3u; k++) st.set_u(SU_GUIDE_CELL0 + k, path, kNoneU);
                if constexpr (SPECTRAL)
                    st.set_v(SV_LAMBDA, path, sample_wavelengths(sample4(p.seed, pixel, who.sample, GROUP_WAVELENGTH).x));
                else
                    st.set_v(SV_LAMBDA, path, spec(0.0f));
                push = true;
            }
        }
        staged.push(0, path, push);
        staged.sync(p, row_ends(plan, n, total, g, r));
    }
}

// ---------------------------------------------------------- light emission

template <bool SPECTRAL> __global__ __launch_bounds__(256) void light_emit_kernel(const Params *plist, uint32_t n) {
    __shared__ RowPlan plan;
    uint32_t total = plan.make(plist, n, [](const Params &q) { return q.pass->light_count * q.pass->sub_samples; });
    __shared__ StagedPushes<1, 256 * kStageIterations> staged;
    staged.init({Q_CLOSEST});
    FOR_ROW_ITEMS(plan, n, total, g, r, i, active) {
        const Params &p = plist[r];
        State st(p);
        const PassState &ps = *p.pass;
        bool push = false;
        uint32_t path = pass_cameras(ps) + i;
        if (active) {
            SlotIdentity who = slot_identity(ps, path);
            uint32_t id = 0x80000000u | who.slot;
            float4 u_point = sample4(p.seed, id, who.sample, GROUP_CAMERA);
            float4 u_pick = sample4(p.seed, id, who.sample, GROUP_WAVELENGTH);
            Spec lambda = SPECTRAL ? sample_wavelengths(u_pick.x) : spec(0.0f);
            float afar = K(p, K_EMIT).x;
            if (u_pick.y < afar) {
                float share;
                uint32_t light = pick_distant(p, u_pick.y / afar, share);
                LightSample s;
                if (light != kNoneU && share > 0.0f && emits_from_afar(p, light) &&
                    sample_light(p, light, V3{0, 0, 0}, u_point.x, u_point.y, s)) {
                    V3 w = s.direction;
                    float radius = K(p, K_EMIT).y;
                    V3 t, b;
                    basis(w, t, b);
                    float2 disk = concentric(u_point.z, u_point.w);
                    V3 start = v3(K(p, K_EMIT_CENTER)) + w * (2.0f * radius) + t * (disk.x * radius) + b * (disk.y * radius);
                    float chance = afar * share * s.pdf;
                    Spec emitted = light_type(p, light) == LIGHT_DOME ? dome_radiance<SPECTRAL>(p, light, w, lambda)
                                                                      : light_radiance<SPECTRAL>(p, light, lambda);
                    Spec carried = emitted * (kPi * radius * radius / chance);
                    float most = spec_max<SPECTRAL>(carried);
                    if (most > 0.0f) {
                        st.set_u(SU_PIXEL, path, kAfarEmitter | light);
                        st.set_u(SU_BOUNCE, path, 0u);
                        st.set_v(SV_ORIGIN, path, f4(start, 0.0f));
                        st.set_v(SV_DIRECTION, path, f4(-w, 0.0f));
                        st.set_v(SV_THROUGHPUT, path, carried);
                        st.set_v(SV_LAMBDA, path, lambda);
                        st.set_v(SV_VERTEX, path, f4(start, 1.0f / most));
                        st.set_v(SV_MIS, path, make_float4(1.0f, chance, 1.0f, -1.0f));
                        push = true;
                    }
                }
            } else {
                uint32_t emitter;
                float chance;
                if (emitter_by_power(p, (u_pick.y - afar) / fmaxf(1e-7f, 1.0f - afar), emitter, chance)) {
                    chance *= 1.0f - afar;
                    uint32_t code = emitter_code(p, emitter);
                    V3 q, n;
                    Spec radiance;
                    float area;
                    bool two_sided = false;
                    if (code & kEmitterTriangle) {
                        uint32_t tri = code & ~kEmitterTriangle;
                        V3 a, e1, e2;
                        triangle_edges(p, tri, a, e1, e2);
                        V3 b = a + e1, c = a + e2;
                        float r = sqrtf(u_point.x);
                        q = a * (1.0f - r) + b * (r * (1.0f - u_point.y)) + c * (r * u_point.y);
                        V3 cab = cross(e1, e2);
                        area = 0.5f * length(cab);
                        n = cab / fmaxf(2.0f * area, 1e-30f);
                        if (u_pick.z < 0.5f) n = -n;
                        two_sided = true;
                        radiance = material_color<SPECTRAL>(p, p.shading[tri].w, 6u, lambda);
                    } else {
                        uint32_t type = light_type(p, code);
                        float4 r0 = light_at(p, code, 0);
                        if (type == LIGHT_SPHERE) {
                            float r = light_at(p, code, 1).w;
                            float z = 1.0f - 2.0f * u_point.x;
                            float s = sqrtf(fmaxf(0.0f, 1.0f - z * z));
                            float phi = 2.0f * kPi * u_point.y;
                            n = V3{s * cosf(phi), s * sinf(phi), z};
                            q = v3(r0) + n * r;
                            area = 4.0f * kPi * r * r;
                        } else {
                            float ax, ay;
                            if (type == LIGHT_RECT) {
                                ax = 2.0f * u_point.x - 1.0f;
                                ay = 2.0f * u_point.y - 1.0f;
                            } else {
                                float2 c = concentric(u_point.x, u_point.y);
                                ax = c.x;
                                ay = c.y;
                            }
                            q = v3(r0) + v3(light_at(p, code, 1)) * ax + v3(light_at(p, code, 2)) * ay;
                            n = v3(light_at(p, code, 3));
                            area = light_at(p, code, 3).w;
                        }
                        radiance = light_radiance<SPECTRAL>(p, code, lambda);
                    }
                    if (area > 0.0f) {
                        float2 disk = concentric(u_point.z, u_point.w);
                        float cos_n = sqrtf(fmaxf(0.0f, 1.0f - (disk.x * disk.x + disk.y * disk.y)));
                        if (cos_n > 0.0f) {
                            V3 t, b;
                            basis(n, t, b);
                            V3 direction = normalize(t * disk.x + b * disk.y + n * cos_n);
                            float point_pdf = chance / area;
                            float direction_pdf = cos_n / (two_sided ? 2.0f * kPi : kPi);
                            Spec throughput = radiance * (cos_n / (point_pdf * direction_pdf));
                            float largest = spec_max<SPECTRAL>(throughput);
                            if (largest > 0.0f) {
                                st.set_u(SU_PIXEL, path, emitter);
                                st.set_u(SU_BOUNCE, path, 0u);
                                st.set_v(SV_ORIGIN, path, f4(offset_origin(q, n), 0.0f));
                                st.set_v(SV_DIRECTION, path, f4(direction, direction_pdf));
                                st.set_v(SV_THROUGHPUT, path, throughput);
                                st.set_v(SV_LAMBDA, path, lambda);
                                st.set_v(SV_VERTEX, path, f4(q, 1.0f / largest));
                                st.set_v(SV_MIS, path, make_float4(1.0f, point_pdf, cos_n, -1.0f));
                                push = true;
                            }
                        }
                    }
                }
            }
        }
        staged.push(0, path, push);
        staged.sync(p, row_ends(plan, n, total, g, r));
    }
}

// ---------------------------------------------------------- closest hits

// The closest hits of one bounce's queue (intersect_closest.comp:45-87): the
// BVH walk, then the analytic lights that the ray meets first. Results go to
// SV_HIT / SU_HIT (kNoneU: a miss); route_kernel then sorts the paths into
// the miss, surface and light queues, so this kernel needs no LDS beyond the
// traversal stack and keeps its occupancy. One ray per thread: every
// persistent schedule measured slower here, its refills stalling on the
// queue-then-state loads. (Testing the lights in route_kernel instead
// measured slower: the state is read twice.)
template <int W, int S, int G>
__global__ __launch_bounds__(G) void closest_kernel_rt(const Params *plist, uint32_t n, uint32_t current) {
    __shared__ uint32_t lds[S * G];
    __shared__ RowPlan plan;
    const uint32_t queue = Q_CLOSEST + current;
    auto count_of = [queue](const Params &q) { return q.queues[queue]; };
    if (idle_block(plist, n, count_of)) return;
    uint32_t total = plan.make(plist, n, count_of);
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            plist[i].queues[Q_SHADOW] = 0u;  // consumed by the last bounce's shadow kernel
            plist[i].pass->closest_rays += plan.count[i];
        }
    FOR_ROW_ITEMS(plan, n, total, g, r, at, active) {
        if (!active) continue;
        const Params &p = plist[r];
        State st(p);
        // The spill column is the item's index in its render's queue: below
        // the render's slots, for which its spill area is laid out.
        const DeviceScene sc = p.scene;  // a copy: registers, not reloads after every store
        uint32_t nlights = light_count(p);
        uint32_t cameras = pass_cameras(*p.pass);
        uint32_t path = queue_path(p, queue, at);
        float4 o4 = st.v(SV_ORIGIN, path), d4 = st.v(SV_DIRECTION, path);
        V3 o = v3(o4), d = v3(d4);
        Hit h = trace_closest<W, S>(sc, o.x, o.y, o.z, d.x, d.y, d.z, kInf, lds + threadIdx.x, G, at);
        uint32_t what = h.slot;
        bool light_path = path >= cameras;
        bool camera_ray = bounce_total(st.u(SU_BOUNCE, path)) == 0u;
        if (nlights) nearest_light(p, o, d, camera_ray, light_path, h.t, what);
        if (what != kNoneU) st.set_v(SV_HIT, path, make_float4(h.t, 0.0f, h.u, h.v));
        st.set_u(SU_HIT, path, what);
    }
}

// The paths of this bounce's closest queue to the miss, surface (camera
// paths) and light (light paths) queues, by their hits.
__global__ __launch_bounds__(256) void route_kernel(const Params *plist, uint32_t n, uint32_t current) {
    __shared__ RowPlan plan;
    const uint32_t queue = Q_CLOSEST + current;
    auto count_of = [queue](const Params &q) { return q.queues[queue]; };
    if (idle_block(plist, n, count_of)) return;
    uint32_t total = plan.make(plist, n, count_of);
    __shared__ StagedPushes<3, 256 * kStageIterations> staged;
    staged.init({Q_MISS, Q_SURFACE, Q_LIGHT});
    FOR_ROW_ITEMS(plan, n, total, g, r, at, active) {
        const Params &p = plist[r];
        State st(p);
        uint32_t cameras = pass_cameras(*p.pass);
        uint32_t path = active ? queue_path(p, queue, at) : 0u;
        uint32_t what = active ? st.u(SU_HIT, path) : kNoneU;
        bool light_path = path >= cameras;
        staged.push(0, path, active && what == kNoneU);
        staged.push(1, path, active && what != kNoneU && !light_path);
        staged.push(2, path, active && what != kNoneU && light_path);
        staged.sync(p, row_ends(plan, n, total, g, r));
    }
}

// Gaussian splats on this bounce's segments (intersect_closest.comp:27-42,
// 65-68), between the closest hits and the routing: a camera path gathers the
// splats' radiance before its hit (or to infinity on a miss) and keeps their
// transmittance of its throughput; a light path is only attenuated. A path
// the splats stop (Russian roulette) becomes a miss with no throughput left.
template <bool SPECTRAL>
__global__ __launch_bounds__(128) void splat_kernel(const Params *plist, uint32_t n, uint32_t current) {
    __shared__ uint32_t stack[kSplatStack * 128];
    __shared__ RowPlan plan;
    const uint32_t queue = Q_CLOSEST + current;
    auto count_of = [queue](const Params &q) { return q.queues[queue]; };
    if (idle_block(plist, n, count_of)) return;
    uint32_t total = plan.make(plist, n, count_of);
    FOR_ROW_ITEMS(plan, n, total, g, r, at, active) {
        if (!active) continue;
        const Params &p = plist[r];
        if (!splats_present(p)) continue;
        State st(p);
        const PassState &ps = *p.pass;
        uint32_t path = queue_path(p, queue, at);
        uint32_t what = st.u(SU_HIT, path);
        float reach = what == kNoneU ? kInf : st.v(SV_HIT, path).x;
        V3 o = v3(st.v(SV_ORIGIN, path)), d = v3(st.v(SV_DIRECTION, path));
        uint32_t bounces = st.u(SU_BOUNCE, path);
        SlotIdentity who = slot_identity(ps, path);
        uint32_t seed = splat_seed(p, who.sample, who.slot, bounces, 0u);
        Spec throughput = st.v(SV_THROUGHPUT, path);
        float kept;
        if (path >= pass_cameras(ps)) {
            kept = splat_transmittance(p, o, d, 0.0f, reach, seed, stack + threadIdx.x, 128u);
        } else {
            Spec added;
            splat_composite<SPECTRAL>(p, o, d, reach, st.v(SV_LAMBDA, path), seed, added, kept, stack + threadIdx.x,
                                      128u);
            st.set_v(SV_RADIANCE, path,
                     st.v(SV_RADIANCE, path) + clamped<SPECTRAL>(p, throughput * added, bounce_total(bounces)));
        }
        if (kept == 1.0f) continue;
        st.set_v(SV_THROUGHPUT, path, throughput * kept);
        if (!(kept > 0.0f)) st.set_u(SU_HIT, path, kNoneU);
    }
}

// The splats' transmittance of this bounce's shadow rays
// (intersect_shadow.comp:61-67), before the shadow kernel walks them: it
// scales what they deliver, in any order with the cut-outs' share.
__global__ __launch_bounds__(128) void splat_shadow_kernel(const Params *plist, uint32_t n) {
    __shared__ uint32_t stack[kSplatStack * 128];
    __shared__ RowPlan plan;
    auto count_of = [](const Params &q) { return q.queues[Q_SHADOW]; };
    if (idle_block(plist, n, count_of)) return;
    uint32_t total = plan.make(plist, n, count_of);
    FOR_ROW_ITEMS(plan, n, total, g, r, at, active) {
        if (!active) continue;
        const Params &p = plist[r];
        if (!splats_present(p)) continue;
        State st(p);
        uint32_t path = queue_path(p, Q_SHADOW, at);
        float4 o = st.v(SV_SHADOW_ORIGIN, path);
        V3 d = v3(st.v(SV_SHADOW_DIRECTION, path));
        SlotIdentity who = slot_identity(*p.pass, path);
        uint32_t seed = splat_seed(p, who.sample, who.slot, st.u(SU_BOUNCE, path), 0x68e31da4u);
        float kept = splat_transmittance(p, v3(o), d, 0.0f, o.w, seed, stack + threadIdx.x, 128u);
        if (kept != 1.0f) st.set_v(SV_SHADOW_RADIANCE, path, st.v(SV_SHADOW_RADIANCE, path) * kept);
    }
}

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

// shade_surface.comp:48-312. PROGRAMS: some material has a node-graph
// program (shader_program.glsl), run once a hit before the surface is read.
template <bool SPECTRAL, bool PROGRAMS>
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
        // The learned guide takes the diffuse lobe's share of the BSDF's
        // choice (its mixtures model incident light, which a glossy lobe's
        // own sampling already follows better); the grid a fixed share
        // (loading.lucb:281).
        if (suited && ready) guided = K(p, K_GUIDE).z * (p.learned ? surface.p_diffuse : 1.0f);
    }

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
            if (guided > 0.0f && !tracing) bsdf_pdf = guided * guide_pdf(p, cell, s.direction) + (1.0f - guided) * bsdf_pdf;
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
    if (u_opacity.y < guided) {
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
    if (guide_on && total < GUIDE_RING) {
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
template <bool SPECTRAL, bool LIGHT, bool PROGRAMS>
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
            else pu = shade_camera<SPECTRAL, PROGRAMS>(p, ps, path);
        }
        staged.push(0, path, pu.closest);
        staged.push(1, path, pu.shadow);
        staged.sync(p, row_ends(plan, n, total, g, r));
    }
}

// ---------------------------------------------------------------- misses

template <bool SPECTRAL> __global__ __launch_bounds__(256) void miss_kernel(const Params *plist, uint32_t n) {
    __shared__ RowPlan plan;
    auto count_of = [](const Params &q) { return q.queues[Q_MISS]; };
    if (idle_block(plist, n, count_of)) return;
    uint32_t total = plan.make(plist, n, count_of);
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            plist[i].queues[Q_SURFACE] = 0u;  // consumed by this bounce's shading
            plist[i].queues[Q_LIGHT] = 0u;
        }
    FOR_ROW_ITEMS(plan, n, total, g, r, at, active) {
        if (!active) continue;
        const Params &p = plist[r];
        State st(p);
        const PassState &ps = *p.pass;
        uint32_t nlights = light_count(p);
        uint32_t path = queue_path(p, Q_MISS, at);
        if (path >= pass_cameras(ps)) continue;  // light paths that leave just end
        uint32_t bounce = bounce_total(st.u(SU_BOUNCE, path));
        float4 ray = st.v(SV_DIRECTION, path);
        V3 rd = v3(ray);
        Spec throughput = st.v(SV_THROUGHPUT, path);
        Spec lambda = st.v(SV_LAMBDA, path);
        Spec radiance = spec_of<SPECTRAL>(K(p, K_BACKGROUND_FIT), K(p, K_BACKGROUND_RGB), lambda);
        for (uint32_t l = 0; l < nlights; l++) {
            uint32_t type = light_type(p, l);
            if (type < LIGHT_DISTANT) continue;
            if (type == LIGHT_DISTANT) {
                float cos_half = light_at(p, l, 2).w;
                if (cos_half >= 1.0f - 1e-7f || dot(rd, -v3(light_at(p, l, 3))) < cos_half) continue;
            }
            float weight = 1.0f;
            if (bounce == 0u) {
                if (!light_visible(p, l)) continue;
            } else if (ray.w > 0.0f) {
                float traced = 0.0f;
                if (light_tracing(p)) {
                    float4 mis = st.v(SV_MIS, path);
                    if (emits_from_afar(p, l) && mis.y > 0.0f)
                        traced = mis.x * (afar_first_density(p, v3(st.v(SV_VERTEX, path)), rd, mis.z) / mis.y) *
                                 (afar_emission_pdf(p, l, rd) / ray.w);
                }
                weight = mis_weight(light_hit_pdf(p, l, V3{0, 0, 0}, rd, kInf) * light_select_pdf(p, l) / ray.w, traced);
            }
            Spec e = type == LIGHT_DOME ? dome_radiance<SPECTRAL>(p, l, rd, lambda) : light_radiance<SPECTRAL>(p, l, lambda);
            radiance = radiance + e * weight;
        }
        Spec contribution = throughput * radiance;
        float clamp_value = K(p, K_INFO).z;
        float largest = spec_max<SPECTRAL>(contribution);
        if (bounce > 0u && clamp_value > 0.0f && largest > clamp_value) contribution = contribution * (clamp_value / largest);
        st.set_v(SV_RADIANCE, path, st.v(SV_RADIANCE, path) + contribution);
    }
}

// ------------------------------------------------------------ shadow rays

// Shadow rays (intersect_shadow.comp:51-92): any hit to the light; with
// cut-outs present, a blocked ray walks through the surfaces it meets (at most
// 16), keeping 1 - opacity of each; an unblocked ray delivers its radiance to
// the camera path or, for a light path, splats it onto the film.
// CUTOUTS false (no material of the scene has geometry opacity below one)
// leaves the walk out, and with it a second traversal's registers.
template <int W, int S, int G, bool CUTOUTS>
__global__ __launch_bounds__(G) void shadow_kernel(const Params *plist, uint32_t n) {
    __shared__ uint32_t lds[S * G];
    __shared__ RowPlan plan;
    auto count_of = [](const Params &q) { return q.queues[Q_SHADOW]; };
    if (idle_block(plist, n, count_of)) return;
    uint32_t total = plan.make(plist, n, count_of);
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            plist[i].queues[Q_MISS] = 0u;  // consumed by this bounce's miss kernel
            plist[i].pass->shadow_rays += plan.count[i];
        }
    FOR_ROW_ITEMS(plan, n, total, g, r, at, active) {
        if (!active) continue;
        const Params &p = plist[r];
        State st(p);
        const PassState &ps = *p.pass;
        // spill column: the item's index in its render's queue (a register copy of the scene)
        const DeviceScene sc = p.scene;
        uint32_t nlights = light_count(p);
        bool cutouts = CUTOUTS && K(p, K_INFO).w >= 0.5f;
        uint32_t path = queue_path(p, Q_SHADOW, at);
        float4 o = st.v(SV_SHADOW_ORIGIN, path);
        V3 d = v3(st.v(SV_SHADOW_DIRECTION, path));
        bool light_path = path >= pass_cameras(ps);
        if (light_path && nlights && light_in_the_way(p, v3(o), d, o.w)) continue;
        uint32_t thread = at;
        bool blocked = trace_any<W, S>(sc, o.x, o.y, o.z, d.x, d.y, d.z, o.w, lds + threadIdx.x, G, thread);
        if (blocked && !cutouts) continue;
        float kept = 1.0f;
        bool deliver = !blocked;
        if (CUTOUTS && blocked) {
            V3 from = v3(o);
            float remaining = o.w;
            for (uint32_t hop = 0; hop < 16u; hop++) {
                Hit h = trace_closest<W, S>(sc, from.x, from.y, from.z, d.x, d.y, d.z, remaining, lds + threadIdx.x, G,
                                            thread);
                if (h.slot == kNone) {
                    deliver = true;
                    break;
                }
                float opacity = clampf(material_at(p, p.shading[h.slot].w, 12).x, 0.0f, 1.0f);
                kept *= 1.0f - opacity;
                if (kept < 1e-4f) break;
                V3 q = from + d * h.t;
                float margin = 1e-4f * fmaxf(1.0f, fmaxf(fabsf(q.x), fmaxf(fabsf(q.y), fabsf(q.z))));
                from = q + d * margin;
                remaining -= h.t + margin;
                if (remaining <= 0.0f) {
                    deliver = true;
                    break;
                }
            }
        }
        if (!deliver) continue;
        float4 value = st.v(SV_SHADOW_RADIANCE, path) * kept;
        if (!light_path) {
            st.set_v(SV_RADIANCE, path, st.v(SV_RADIANCE, path) + value);
        } else {
            if (isnan(value.x) || isnan(value.y) || isnan(value.z) || isinf(value.x) || isinf(value.y) || isinf(value.z))
                continue;
            float *f = reinterpret_cast<float *>(&p.film[(size_t)FILM_SPLAT * p.pixels + st.u(SU_SPLAT, path)]);
            atomicAdd(&f[0], value.x);
            atomicAdd(&f[1], value.y);
            atomicAdd(&f[2], value.z);
        }
    }
}

// ------------------------------------------------------------- the film

// One thread a camera path of the band; a pass of several samples adds that
// path's samples to its pixel one after another, in sample order, as that
// many passes of one sample would.
template <bool SPECTRAL> __global__ __launch_bounds__(256) void finish_kernel(const Params *plist, uint32_t n) {
    __shared__ RowPlan plan;
    uint32_t total = plan.make(plist, n, [](const Params &q) { return q.pass->band_count; });
    FOR_ROW_ITEMS(plan, n, total, g, r, item, active) {
        if (!active) continue;
        const Params &p = plist[r];
        State st(p);
        const PassState &ps = *p.pass;
        bool guide_on = guiding(p);
        for (uint32_t j = 0; j < ps.sub_samples; j++) {
            uint32_t path = j * ps.band_count + item;
            uint32_t sample = ps.sample + j;
            uint32_t pixel = st.u(SU_PIXEL, path);
            if (pixel == kNoneU) continue;
            Spec radiance = st.v(SV_RADIANCE, path), lambda = st.v(SV_LAMBDA, path);
            V3 value = spec_to_film<SPECTRAL>(p, radiance, lambda);
            if (isnan(value.x) || isnan(value.y) || isnan(value.z) || isinf(value.x) || isinf(value.y) || isinf(value.z))
                value = V3{0, 0, 0};
            float4 &sum = p.film[pixel];
            float4 s = sum;
            sum = make_float4(s.x + value.x, s.y + value.y, s.z + value.z, s.w + 1.0f);
            if (sample & 1u) {
                float4 &half = p.film[(size_t)p.pixels + pixel];
                float4 h = half;
                half = make_float4(h.x + 2.0f * value.x, h.y + 2.0f * value.y, h.z + 2.0f * value.z, h.w);
            }
            if (guide_on && !p.learned && hash_u32(pixel * 9781u + sample) % GUIDE_TRAINING == 0u) {
                float found = spec_luminance<SPECTRAL>(p, radiance, lambda);
                for (uint32_t k = 0; k < GUIDE_RING; k++) {
                    uint32_t cell = st.u(SU_GUIDE_CELL0 + k, path);
                    if (cell == kNoneU) break;
                    float4 v = st.v(SV_GUIDE0 + k, path);
                    if (v.y <= 0.0f) continue;
                    V3 dir = octahedral_decode(st.u(SU_GUIDE_DIR0 + k, path));
                    // The light found past the vertex (finish.comp's found - v.z).
                    // Where nothing came, the two f32 sums differ only by their
                    // rounding; such a difference is no sample (it would count,
                    // and fit a lobe of kappa 1000 to one stray direction).
                    float gain = found - v.z;
                    if (gain <= 1e-5f * fmaxf(found, v.z)) gain = 0.0f;
                    guide_train(p, cell, dir, gain / v.y * v.w, v.x);
                }
            }
        }
    }
}

__global__ __launch_bounds__(256) void guide_fit_kernel(const Params *plist) {
    const Params &p = plist[blockIdx.y];
    if (!p.guide) return;
    for (uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x; cell < GUIDE_CELLS; cell += gridDim.x * blockDim.x) {
        if (guide_keys(p)[cell] == 0u) continue;
        float *s = guide_stats(p);
        float4 *l = guide_lobes(p);
        float w4[4];
        for (uint32_t k = 0; k < GUIDE_LOBES; k++) {
            float *at = s + (size_t)(cell * GUIDE_LOBES + k) * 5u;
            float w = at[0];
            V3 sum{at[1], at[2], at[3]};
            if (w > 0.0f && dot(sum, sum) > 0.0f) {
                float r = clampf(length(sum) / w, 0.0f, 0.999f);
                float kappa = clampf(r * (3.0f - r * r) / (1.0f - r * r), 0.1f, 1000.0f);
                l[cell * (GUIDE_LOBES + 1u) + k] = f4(normalize(sum), kappa);
            }
            w4[k] = w;
            for (uint32_t c = 0; c < 5u; c++) at[c] *= 0.9f;
        }
        if (w4[0] + w4[1] + w4[2] + w4[3] > 0.0f) l[cell * (GUIDE_LOBES + 1u) + GUIDE_LOBES] = make_float4(w4[0], w4[1], w4[2], w4[3]);
    }
}

// Adaptive sampling's convergence check (adaptive.comp:14-40; Dammertz et
// al.): phase 0 gives each pixel its own verdict, |full - half| / (1e-4 +
// sqrt(brightness)) below the noise threshold, in the scratch plane; phase 1
// marks a pixel converged (FILM_HALF.w = 1, which the camera kernel skips)
// when its whole 3x3 neighbourhood is. Runs after the last band of every 4th
// sample from the adaptive minimum on (integrator.lucb:818-826).
__device__ __forceinline__ bool adaptive_now(const Params &p) {
    const PassState &ps = *p.pass;
    uint32_t taken = ps.sample + ps.sub_samples;  // passes with a noise threshold carry one sample
    return K(p, K_ADAPTIVE).x > 0.0f && ps.band + 1u == ps.bands && (float)taken >= K(p, K_ADAPTIVE).y &&
           taken % 4u == 0u;
}

__global__ __launch_bounds__(256) void adaptive_kernel(const Params *plist, uint32_t n, uint32_t phase) {
    __shared__ RowPlan plan;
    uint32_t total = plan.make(plist, n, [](const Params &q) { return adaptive_now(q) ? q.pixels : 0u; });
    FOR_ROW_ITEMS(plan, n, total, g, r, pixel, active) {
        if (!active) continue;
        const Params &p = plist[r];
        if (phase == 0u) {
            float4 sum = p.film[pixel], half = p.film[(size_t)FILM_HALF * p.pixels + pixel];
            float scale = 1.0f / fmaxf(sum.w, 1.0f);
            float difference = (fabsf(sum.x - half.x) + fabsf(sum.y - half.y) + fabsf(sum.z - half.z)) * scale;
            float brightness = fmaxf(0.0f, (sum.x + sum.y + sum.z) * scale);
            float error = difference / (0.0001f + sqrtf(brightness));
            p.film[(size_t)FILM_SCRATCH * p.pixels + pixel] =
                make_float4(error < K(p, K_ADAPTIVE).x ? 1.0f : 0.0f, 0.0f, 0.0f, 0.0f);
        } else {
            int w = (int)p.width, h = (int)p.height;
            int x = (int)(pixel % p.width), y = (int)(pixel / p.width);
            float converged = 1.0f;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int nx = x + dx, ny = y + dy;
                    if (nx < 0 || ny < 0 || nx >= w || ny >= h) continue;
                    converged = fminf(converged, p.film[(size_t)FILM_SCRATCH * p.pixels + (size_t)ny * w + nx].x);
                }
            float4 &half = p.film[(size_t)FILM_HALF * p.pixels + pixel];
            half.w = converged;
        }
    }
}

// Pixels marked converged, added into *out.
__global__ __launch_bounds__(256) void converged_kernel(const Params *plist, uint32_t *out) {
    const Params &p = plist[0];
    uint32_t count = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.pixels; i += gridDim.x * blockDim.x)
        count += p.film[(size_t)FILM_HALF * p.pixels + i].w > 0.5f ? 1u : 0u;
    for (uint32_t off = 32u; off > 0; off >>= 1) count += __shfl_xor(count, (int)off);
    if ((threadIdx.x & 63u) == 0u && count) atomicAdd(out, count);
}

// Khronos PBR Neutral (film_convert.comp:21-34).
__device__ __forceinline__ V3 neutral(V3 c) {
    const float start = 0.8f - 0.04f, desaturation = 0.15f;
    float x = fminf(c.x, fminf(c.y, c.z));
    float offset = x < 0.08f ? x - 6.25f * x * x : 0.04f;
    c = c - V3{offset, offset, offset};
    float peak = fmaxf(c.x, fmaxf(c.y, c.z));
    if (peak < start) return c;
    float d = 1.0f - start;
    float new_peak = 1.0f - d * d / (peak + d - start);
    c = c * (new_peak / peak);
    float g = 1.0f - 1.0f / (desaturation * (peak - new_peak) + 1.0f);
    return c * (1.0f - g) + V3{new_peak, new_peak, new_peak} * g;
}

// Only the sample's last band converts (integrator.lucb:672-676).
template <bool SPECTRAL>
__global__ __launch_bounds__(256) void film_convert_kernel(const Params *plist, uint32_t n, uint32_t always) {
    __shared__ RowPlan plan;
    uint32_t total = plan.make(plist, n, [always](const Params &q) {
        return always || q.pass->band + 1u == q.pass->bands ? q.pixels : 0u;
    });
    FOR_ROW_ITEMS(plan, n, total, g, r, pixel, active) {
        if (!active) continue;
        const Params &p = plist[r];
        uint32_t samples = p.pass->light_samples;
        float4 sum = p.film[pixel];
        V3 value = v3(sum) / fmaxf(sum.w, 1.0f);
        if (samples > 0u) value = value + v3(p.film[(size_t)FILM_SPLAT * p.pixels + pixel]) / (float)samples;
        if constexpr (SPECTRAL)
            value = V3{dot(v3(K(p, K_FILM)), value), dot(v3(K(p, K_FILM + 1)), value), dot(v3(K(p, K_FILM + 2)), value)};
        p.image[pixel] = f4(value, 1.0f);
        float e = exp2f(K(p, K_DISPLAY).x);
        V3 v = value * e;
        V3 shown{1.70505f * v.x - 0.62179f * v.y - 0.08326f * v.z, -0.13026f * v.x + 1.14080f * v.y - 0.01055f * v.z,
                 -0.02400f * v.x - 0.12897f * v.y + 1.15297f * v.z};
        shown = V3{fmaxf(shown.x, 0.0f), fmaxf(shown.y, 0.0f), fmaxf(shown.z, 0.0f)};
        if (K(p, K_DISPLAY).y > 0.5f) shown = neutral(shown);
        p.display[pixel] = make_float4(fminf(shown.x, 1.0f), fminf(shown.y, 1.0f), fminf(shown.z, 1.0f), 1.0f);
    }
}

// ------------------------------------------------------------ albedo tables
// tables.comp: F · G2 / G1 averaged over a 32x32 stratified grid of visible
// normals per node.

__device__ float table_albedo(float mu, float roughness, float eta) {
    V3 wo{sqrtf(1.0f - mu * mu), 0.0f, mu};
    float2 a = ggx_alpha(roughness, 0.0f);
    float sum = 0.0f;
    for (uint32_t i = 0; i < 32u; i++)
        for (uint32_t j = 0; j < 32u; j++) {
            V3 h = ggx_sample(wo, a, ((float)i + 0.5f) / 32.0f, ((float)j + 0.5f) / 32.0f);
            V3 wi = reflect(-wo, h);
            if (wi.z <= 0.0f) continue;
            float f = eta > 0.0f ? fresnel_dielectric(dot(wo, h), eta) : 1.0f;
            sum += f * ggx_g2(wi, wo, a) / ggx_g1(wo, a);
        }
    return sum / 1024.0f;
}

__global__ __launch_bounds__(64) void tables_kernel(float *t, uint32_t phase) {
    uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (phase == 0u) {
        uint32_t plane = TABLE_MU * TABLE_ROUGHNESS;
        if (cell >= plane * (1u + TABLE_ETA)) return;
        uint32_t layer = cell / plane, at = cell % plane;
        float mu = table_mu(at % TABLE_MU), r = table_roughness(at / TABLE_MU);
        if (layer == 0u) t[TABLE_GGX + at] = table_albedo(mu, r, -1.0f);
        else t[TABLE_DIELECTRIC + (layer - 1u) * plane + at] = table_albedo(mu, r, eta_of_z(table_z(layer - 1u)));
        return;
    }
    if (cell >= TABLE_ROUGHNESS) return;
    float sum = 0.0f;
    for (uint32_t m = 0; m + 1u < TABLE_MU; m++) {
        float a = table_mu(m), b = table_mu(m + 1u);
        sum += 0.5f * (t[TABLE_GGX + cell * TABLE_MU + m] * a + t[TABLE_GGX + cell * TABLE_MU + m + 1u] * b) * (b - a);
    }
    t[TABLE_GGX_AVERAGE + cell] = 2.0f * sum;
}

// A fill kernel (graph-safe in place of memset nodes).
// Bounding spheres of the analytic lights (see nearest_light): group g's
// sphere at out[g], light l's at out[groups + l]. One thread per group.
__global__ __launch_bounds__(64) void light_bounds_kernel(const float4 *lights, uint32_t n, float4 *out) {
    uint32_t groups = (n + kLightGroup - 1u) / kLightGroup;
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= groups) return;
    uint32_t first = g * kLightGroup, end = min(n, first + kLightGroup);
    float4 sph[kLightGroup];
    V3 centre{0, 0, 0};
    uint32_t hittable = 0;
    for (uint32_t l = first; l < end; l++) {
        const float4 *r = lights + (size_t)l * 7u;
        uint32_t type = (uint32_t)r[0].w;
        float radius = -1.0f;
        if (type == LIGHT_SPHERE) radius = r[1].w;
        else if (type < LIGHT_DISTANT) radius = length(v3(r[1])) + length(v3(r[2]));
        V3 c = v3(r[0]);
        if (radius >= 0.0f) {
            float scale = fmaxf(fabsf(c.x), fmaxf(fabsf(c.y), fabsf(c.z)));
            radius = radius * 1.001f + 1e-5f * scale + 1e-30f;
            centre = centre + c;
            hittable++;
        }
        sph[l - first] = f4(c, radius);
        out[groups + l] = sph[l - first];
    }
    float group_radius = -1.0f;
    if (hittable) {
        centre = centre * (1.0f / (float)hittable);
        group_radius = 0.0f;
        for (uint32_t l = first; l < end; l++) {
            float4 s = sph[l - first];
            if (s.w >= 0.0f) group_radius = fmaxf(group_radius, length(v3(s) - centre) + s.w);
        }
        float scale = fmaxf(fabsf(centre.x), fmaxf(fabsf(centre.y), fabsf(centre.z)));
        group_radius = group_radius * 1.001f + 1e-5f * scale + 1e-30f;
    }
    out[g] = f4(centre, group_radius);
}

__global__ __launch_bounds__(256) void fill_kernel(uint32_t *dst, size_t words, uint32_t value) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = value;
}

}  // namespace rt

// ------------------------------------------------------------------ launchers

namespace {

// Workgroups of the pushing stages: about 32 waves a compute unit, so that
// each workgroup loops over several iterations between its staged flushes.
// (The tracing kernels take one item per thread: measured faster than any
// persistent grid, whose slowest workgroups set the launch's end.)
uint32_t staged_blocks(uint32_t group, uint32_t wanted) {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
    }
    uint64_t cap = (uint64_t)cus * 32u / (group / 64u);
    return wanted < cap ? wanted : (uint32_t)cap;
}

template <int W, int S, int G>
void launch_closest(const rt::Params *p, uint32_t n, uint32_t cur, uint32_t blocks, hipStream_t s) {
    hipLaunchKernelGGL((rt::closest_kernel_rt<W, S, G>), dim3(blocks), dim3(G), 0, s, p, n, cur);
}
template <int W, int S, int G> void launch_shadow(const rt::Params *p, uint32_t n, uint32_t blocks, bool cutouts,
                                                  hipStream_t s) {
    if (cutouts) hipLaunchKernelGGL((rt::shadow_kernel<W, S, G, true>), dim3(blocks), dim3(G), 0, s, p, n);
    else hipLaunchKernelGGL((rt::shadow_kernel<W, S, G, false>), dim3(blocks), dim3(G), 0, s, p, n);
}

template <int W, int S, int G> void fill_trace(RenderTraceVariant &v) {
    v.closest = &launch_closest<W, S, G>;
    v.shadow = &launch_shadow<W, S, G>;
    v.group = G;
}
template <int W, int S> void fill_trace_g(RenderTraceVariant *t, int wi, int si) {
    fill_trace<W, S, 64>(t[(wi * kStackSizes + si) * 3 + 0]);
    fill_trace<W, S, 128>(t[(wi * kStackSizes + si) * 3 + 1]);
    fill_trace<W, S, 256>(t[(wi * kStackSizes + si) * 3 + 2]);
}
template <int W> void fill_trace_s(RenderTraceVariant *t, int wi) {
    fill_trace_g<W, 16>(t, wi, 0);
    fill_trace_g<W, 24>(t, wi, 1);
    fill_trace_g<W, 32>(t, wi, 2);
    fill_trace_g<W, 48>(t, wi, 3);
}

}  // namespace

const RenderTraceVariant *render_trace_variants() {
    static RenderTraceVariant table[kTraceVariants];
    static bool ready = false;
    if (!ready) {
        fill_trace_s<2>(table, 0);
        fill_trace_s<4>(table, 1);
        fill_trace_s<8>(table, 2);
        fill_trace_s<kQuantizedLayout + 4>(table, 3);
        fill_trace_s<kQuantizedLayout + 8>(table, 4);
        ready = true;
    }
    return table;
}

template <bool SPECTRAL>
static void launch_stage(Stage stage, const rt::Params *p, uint32_t n, uint32_t arg, uint32_t blocks, hipStream_t s) {
#define LR_STAGE(kernel, group, ...)                                                             \
    hipLaunchKernelGGL(kernel, dim3(staged_blocks(group, blocks)), dim3(group), 0, s, __VA_ARGS__)
    switch (stage) {
        case Stage::Camera: LR_STAGE((rt::camera_kernel<SPECTRAL>), 256, p, n); break;
        case Stage::LightEmit: LR_STAGE((rt::light_emit_kernel<SPECTRAL>), 256, p, n); break;
        case Stage::Shade: LR_STAGE((rt::shade_kernel<SPECTRAL, false, false>), 128, p, n, arg); break;
        case Stage::ShadeLight: LR_STAGE((rt::shade_kernel<SPECTRAL, true, false>), 128, p, n, arg); break;
        case Stage::ShadeProgram: LR_STAGE((rt::shade_kernel<SPECTRAL, false, true>), 128, p, n, arg); break;
        case Stage::ShadeLightProgram: LR_STAGE((rt::shade_kernel<SPECTRAL, true, true>), 128, p, n, arg); break;
        case Stage::Route: LR_STAGE(rt::route_kernel, 256, p, n, arg); break;
        case Stage::Splat: LR_STAGE((rt::splat_kernel<SPECTRAL>), 128, p, n, arg); break;
        case Stage::SplatShadow: LR_STAGE(rt::splat_shadow_kernel, 128, p, n); break;
        case Stage::Miss: LR_STAGE((rt::miss_kernel<SPECTRAL>), 256, p, n); break;
        case Stage::Finish: LR_STAGE((rt::finish_kernel<SPECTRAL>), 256, p, n); break;
        case Stage::Convert: LR_STAGE((rt::film_convert_kernel<SPECTRAL>), 256, p, n, arg); break;
        case Stage::Adaptive: LR_STAGE(rt::adaptive_kernel, 256, p, n, arg); break;
        default: break;
    }
#undef LR_STAGE
}

void launch_render_stage(Stage stage, bool spectral, const rt::Params *p, uint32_t n, uint32_t arg, uint32_t blocks,
                         hipStream_t s) {
    if (n == 0) return;
    switch (stage) {
        case Stage::BeginPass: hipLaunchKernelGGL(rt::begin_pass_kernel, dim3(n), dim3(64), 0, s, p); return;
        case Stage::EndPass: hipLaunchKernelGGL(rt::end_pass_kernel, dim3(n), dim3(64), 0, s, p); return;
        case Stage::GuideFit: hipLaunchKernelGGL(rt::guide_fit_kernel, dim3(blocks, n), dim3(256), 0, s, p); return;
        default: break;
    }
    if (spectral) launch_stage<true>(stage, p, n, arg, blocks, s);
    else launch_stage<false>(stage, p, n, arg, blocks, s);
}

void launch_tables(float *tables, hipStream_t s) {
    uint32_t cells = rt::TABLE_MU * rt::TABLE_ROUGHNESS * (1u + rt::TABLE_ETA);
    hipLaunchKernelGGL(rt::tables_kernel, dim3((cells + 63) / 64), dim3(64), 0, s, tables, 0u);
    hipLaunchKernelGGL(rt::tables_kernel, dim3(1), dim3(64), 0, s, tables, 1u);
}

void launch_converged(const rt::Params *p, uint32_t pixels, uint32_t *out, hipStream_t s) {
    uint32_t blocks = (uint32_t)std::min<uint64_t>(((uint64_t)pixels + 255) / 256, 4096);
    hipLaunchKernelGGL(rt::converged_kernel, dim3(blocks ? blocks : 1), dim3(256), 0, s, p, out);
}

void launch_light_bounds(const float4 *lights, uint32_t n, float4 *out, hipStream_t s) {
    uint32_t groups = (n + rt::kLightGroup - 1u) / rt::kLightGroup;
    if (groups == 0) return;
    hipLaunchKernelGGL(rt::light_bounds_kernel, dim3((groups + 63) / 64), dim3(64), 0, s, lights, n, out);
}

void launch_fill(uint32_t *dst, size_t words, uint32_t value, hipStream_t s) {
    if (words == 0) return;
    uint32_t blocks = (uint32_t)std::min<size_t>((words + 255) / 256, 8192);
    hipLaunchKernelGGL(rt::fill_kernel, dim3(blocks), dim3(256), 0, s, dst, words, value);
}

}  // namespace lr
