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

#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "render_bsdf.h"
#include "render_guide.h"
#include "render_program.h"
#include "render_splats.h"
#include "trace_kernels.h"

namespace lr {
namespace rt {

}  // namespace rt
}  // namespace lr

#include "render_connect.h"

namespace lr {
namespace rt {

// ------------------------------------------------------------ pass control


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
        // the learned guide trains after each pass of a quarter of its training samples
        if (p.learned && s->sample < p.train_samples)
            k = min(k, min(p.train_samples - s->sample, max(1u, p.train_samples / 4u)));
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

}  // namespace rt
}  // namespace lr

#include "render_shade.h"

namespace lr {
namespace rt {

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
        uint32_t path = queue_path(p, Q_MISS, at);
        if (path >= pass_cameras(ps)) continue;  // light paths that leave just end
        uint32_t bounce = bounce_total(st.u(SU_BOUNCE, path));
        float4 ray = st.v(SV_DIRECTION, path);
        V3 rd = v3(ray);
        Spec throughput = st.v(SV_THROUGHPUT, path);
        Spec lambda = st.v(SV_LAMBDA, path);
        Spec radiance = spec_of<SPECTRAL>(K(p, K_BACKGROUND_FIT), K(p, K_BACKGROUND_RGB), lambda);
        // the distant and dome lights, in index order (the lights that can
        // be hit add nothing to a miss)
        for (uint32_t k = 0; k < p.far_count; k++) {
            uint32_t l = p.far_lights[k];
            uint32_t type = light_type(p, l);
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

__global__ __launch_bounds__(256) void fill_kernel(uint32_t *dst, size_t words, uint32_t value) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = value;
}

}  // namespace rt

// ------------------------------------------------------------------ launchers

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

namespace {

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

std::vector<rt::LightNode> light_tree(const float *lights, uint32_t n) {
    // Each hittable light's box (a sphere's, or the rectangle / disk's
    // c +- (|u| + |v|) per axis) in f64, grown by 1e-4 of the lights' scale
    // and rounded outward to f32; a light with a non-finite number gets an
    // infinite box (always visited).
    struct Item {
        uint64_t key;
        uint32_t light;
        float lo[3], hi[3];
        double c[3];
    };
    std::vector<Item> items;
    double scale = 1.0;
    for (uint32_t l = 0; l < n; l++) {
        const float *r = lights + (size_t)l * 28u;
        if ((uint32_t)r[3] >= rt::LIGHT_DISTANT) continue;
        for (int a = 0; a < 3; a++) {
            double half = (uint32_t)r[3] == rt::LIGHT_SPHERE ? std::fabs((double)r[7])
                                                             : std::fabs((double)r[4 + a]) + std::fabs((double)r[8 + a]);
            if (std::isfinite(r[a]) && std::isfinite(half)) scale = std::max(scale, std::fabs((double)r[a]) + half);
        }
    }
    double margin = 1e-4 * scale;
    double box_lo[3] = {INFINITY, INFINITY, INFINITY}, box_hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t l = 0; l < n; l++) {
        const float *r = lights + (size_t)l * 28u;
        uint32_t type = (uint32_t)r[3];
        if (type >= rt::LIGHT_DISTANT) continue;
        Item it{0, l, {}, {}, {}};
        for (int a = 0; a < 3; a++) {
            double half = type == rt::LIGHT_SPHERE ? std::fabs((double)r[7])
                                                    : std::fabs((double)r[4 + a]) + std::fabs((double)r[8 + a]);
            double c = r[a], lo = c - half - margin, hi = c + half + margin;
            bool finite = std::isfinite(lo) && std::isfinite(hi);
            it.lo[a] = finite ? std::nextafter((float)lo, -INFINITY) : -INFINITY;
            it.hi[a] = finite ? std::nextafter((float)hi, INFINITY) : INFINITY;
            it.c[a] = finite ? c : 0.0;
            box_lo[a] = std::min(box_lo[a], it.c[a]);
            box_hi[a] = std::max(box_hi[a], it.c[a]);
        }
        items.push_back(it);
    }
    auto spread = [](uint32_t v) {  // 10 bits to every third of 30
        uint64_t x = v & 1023u;
        x = (x | (x << 16)) & 0x030000FFull;
        x = (x | (x << 8)) & 0x0300F00Full;
        x = (x | (x << 4)) & 0x030C30C3ull;
        x = (x | (x << 2)) & 0x09249249ull;
        return x;
    };
    for (Item &it : items)
        for (int a = 0; a < 3; a++) {
            double extent = box_hi[a] - box_lo[a];
            double f = extent > 0.0 ? (it.c[a] - box_lo[a]) / extent : 0.0;
            it.key |= spread((uint32_t)std::min(1023.0, std::max(0.0, f * 1023.0))) << a;
        }
    std::sort(items.begin(), items.end(),
              [](const Item &x, const Item &y) { return x.key != y.key ? x.key < y.key : x.light < y.light; });
    // The Morton order as the leaves of a complete binary tree (padded to a
    // power of two with empty leaves), breadth-first from index 1, each
    // interior box the union of its children's.
    std::vector<rt::LightNode> nodes(light_tree_nodes((uint32_t)items.size()));
    const rt::LightNode empty{{INFINITY, INFINITY, INFINITY}, 0u, {-INFINITY, -INFINITY, -INFINITY}, rt::kNoneU};
    std::fill(nodes.begin(), nodes.end(), empty);
    size_t leaves = nodes.size() / 2;
    for (size_t k = 0; k < items.size(); k++) {
        rt::LightNode &nd = nodes[leaves + k];
        for (int a = 0; a < 3; a++) nd.lo[a] = items[k].lo[a], nd.hi[a] = items[k].hi[a];
        nd.light = items[k].light;
    }
    for (size_t k = leaves; k-- > 1;)
        for (int a = 0; a < 3; a++) {
            nodes[k].lo[a] = std::min(nodes[2 * k].lo[a], nodes[2 * k + 1].lo[a]);
            nodes[k].hi[a] = std::max(nodes[2 * k].hi[a], nodes[2 * k + 1].hi[a]);
        }
    return nodes;
}

uint32_t light_tree_nodes(uint32_t n) {
    uint32_t leaves = 1;
    while (leaves < n) leaves *= 2;
    return n ? 2u * leaves : 0u;
}

std::vector<uint32_t> far_lights(const float *lights, uint32_t n) {
    std::vector<uint32_t> out;
    for (uint32_t l = 0; l < n; l++)
        if ((uint32_t)lights[(size_t)l * 28u + 3u] >= rt::LIGHT_DISTANT) out.push_back(l);
    return out;
}

hipError_t upload_light_tree(const std::vector<rt::LightNode> &nodes, rt::LightNode *out, hipStream_t s) {
    if (nodes.empty()) return hipSuccess;
    hipError_t e = hipMemcpyAsync(out, nodes.data(), nodes.size() * sizeof(rt::LightNode), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);  // `nodes` is the caller's
    return e;
}

void launch_fill(uint32_t *dst, size_t words, uint32_t value, hipStream_t s) {
    if (words == 0) return;
    uint32_t blocks = (uint32_t)std::min<size_t>((words + 255) / 256, 8192);
    hipLaunchKernelGGL(rt::fill_kernel, dim3(blocks), dim3(256), 0, s, dst, words, value);
}

}  // namespace lr
