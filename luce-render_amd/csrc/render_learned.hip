// The learned guide inside the wavefront integrator: each bounce, the network
// gives every camera path's surface hit its mixture (shading then samples and
// weighs it exactly as it does a grid cell's); after each of the first
// samples, Adam steps on batches drawn from the paths' guide records (the
// same records the grid trains on: point, normal, outgoing direction, and
// the radiance found along it over its sampling density).
#include <hip/hip_runtime.h>

#include "guide_net_kernels.h"
#include "render_common.h"
#include "render_guide.h"
#include "render_learned.h"
#include "render_lights.h"

namespace lr {
namespace rt {

namespace {

// The hit point and the geometric normal facing the incoming ray of the
// bounce's surface queue entries (light hits get a placeholder: their
// shading never consults the guide).
struct SurfacePoints {
    const Params &p;
    __device__ __forceinline__ void point(uint32_t i, float *x, float *n) const {
        State st(p);
        uint32_t path = queue_path(p, Q_SURFACE, i);
        float4 hit = st.v(SV_HIT, path);
        V3 o = v3(st.v(SV_ORIGIN, path)), d = v3(st.v(SV_DIRECTION, path));
        uint32_t what = st.u(SU_HIT, path);
        V3 at = o + d * hit.x, g{0.0f, 0.0f, 1.0f};
        if (!(what & kLightHit)) {
            V3 p0, e1, e2;
            triangle_edges(p, what, p0, e1, e2);
            g = normalize(cross(e1, e2));
            if (dot(g, d) > 0.0f) g = -g;
        }
        x[0] = at.x, x[1] = at.y, x[2] = at.z;
        n[0] = g.x, n[1] = g.y, n[2] = g.z;
    }
    __device__ __forceinline__ void store(uint32_t i, const net::Mixture &m) const {
        uint32_t path = queue_path(p, Q_SURFACE, i);
        V3 mu[GUIDE_LOBES];
        for (uint32_t k = 0; k < GUIDE_LOBES; k++) mu[k] = V3{m.mu[k][0], m.mu[k][1], m.mu[k][2]};
        store_learned_mix(p, path, mu, m.kappa, m.w);
    }
};

// The first sample's camera paths (slots 0 .. band_count - 1) at their first
// hit: the first vertex of every sample of a pass takes the mixture of its
// pixel's first sample (its slot modulo band_count), a point within the same
// pixel. A path whose first hit is no surface gets a placeholder point; its
// mixture is still a distribution for the pixel's other samples.
struct FirstSamplePoints {
    const Params &p;
    __device__ __forceinline__ void point(uint32_t path, float *x, float *n) const {
        State st(p);
        uint32_t what = st.u(SU_HIT, path);
        x[0] = x[1] = x[2] = 0.0f;
        n[0] = 0.0f, n[1] = 0.0f, n[2] = 1.0f;
        if (st.u(SU_PIXEL, path) == kNoneU || what == kNoneU || (what & kLightHit)) return;
        float4 hit = st.v(SV_HIT, path);
        V3 o = v3(st.v(SV_ORIGIN, path)), d = v3(st.v(SV_DIRECTION, path));
        V3 at = o + d * hit.x, p0, e1, e2;
        triangle_edges(p, what, p0, e1, e2);
        V3 g = normalize(cross(e1, e2));
        if (dot(g, d) > 0.0f) g = -g;
        x[0] = at.x, x[1] = at.y, x[2] = at.z;
        n[0] = g.x, n[1] = g.y, n[2] = g.z;
    }
    __device__ __forceinline__ void store(uint32_t path, const net::Mixture &m) const {
        V3 mu[GUIDE_LOBES];
        for (uint32_t k = 0; k < GUIDE_LOBES; k++) mu[k] = V3{m.mu[k][0], m.mu[k][1], m.mu[k][2]};
        store_learned_mix(p, path, mu, m.kappa, m.w);
    }
};

// Training samples: record k of path `path`, both drawn by a hash of the
// sample index and the step's salt over the band's camera paths.
template <bool SPECTRAL> struct PathRecords {
    const Params &p;
    uint32_t salt;
    __device__ __forceinline__ net::Sample sample(uint32_t i) const {
        State st(p);
        net::Sample s{};
        s.valid = false;
        uint32_t band = pass_cameras(*p.pass);
        if (band == 0) return s;
        uint32_t h = hash_u32(i * 0x9E3779B9u ^ salt);
        uint32_t path = h % band, k = hash_u32(h ^ 0x85EBCA6Bu) % GUIDE_RING;
        if (st.u(SU_PIXEL, path) == kNoneU || st.u(SU_GUIDE_CELL0 + k, path) == kNoneU) return s;
        float4 v = st.v(SV_GUIDE0 + k, path);  // pdf, throughput luminance, radiance found before, cos
        if (!(v.x > 0.0f) || !(v.y > 0.0f)) return s;
        float found = spec_luminance<SPECTRAL>(p, st.v(SV_RADIANCE, path), st.v(SV_LAMBDA, path));
        float weight = fmaxf(0.0f, found - v.z) / v.y * v.w / v.x;
        if (!(weight > 0.0f) || isinf(weight)) return s;
        float4 at = st.v(SV_GUIDE_POS0 + k, path);
        V3 n = octahedral_decode(__float_as_uint(at.w));
        V3 d = octahedral_decode(st.u(SU_GUIDE_DIR0 + k, path));
        s.p[0] = at.x, s.p[1] = at.y, s.p[2] = at.z;
        s.n[0] = n.x, s.n[1] = n.y, s.n[2] = n.z;
        s.d[0] = d.x, s.d[1] = d.y, s.d[2] = d.z;
        s.weight = weight;
        s.valid = true;
        return s;
    }
};

__device__ __forceinline__ bool training(const Params &p) { return p.learned && p.pass->sample < p.train_samples; }
__device__ __forceinline__ uint32_t salt(const Params &p, uint32_t step) {
    return hash_u32(p.pass->sample * 7919u + step * 104729u + 0x2545F491u);
}

}  // namespace

__global__ __launch_bounds__(64) void learned_infer_kernel(const Params *plist) {
    const Params &p = plist[0];
    if (!net_ready(p)) return;
    SurfacePoints pts{p};
    net::infer_body(p.net, pts, p.queues[Q_SURFACE]);
}

__global__ __launch_bounds__(64) void learned_infer_first_kernel(const Params *plist) {
    const Params &p = plist[0];
    if (!net_ready(p)) return;
    FirstSamplePoints pts{p};
    net::infer_body(p.net, pts, p.pass->band_count);
}

template <bool SPECTRAL> __global__ __launch_bounds__(256) void learned_weights_kernel(const Params *plist, uint32_t step) {
    const Params &p = plist[0];
    if (!training(p)) return;
    net::NetBuffers n = p.net;
    PathRecords<SPECTRAL> src{p, salt(p, step)};
    net::weights_body(n, src, p.train_batch);
}

template <bool SPECTRAL>
__global__ __launch_bounds__(net::kGradThreads) void learned_grad_kernel(const Params *plist, uint32_t step) {
    const Params &p = plist[0];
    if (!training(p)) return;
    net::NetBuffers n = p.net;
    PathRecords<SPECTRAL> src{p, salt(p, step)};
    net::grad_body(n, src, p.train_batch, p.net_partials);
}

__global__ __launch_bounds__(256) void learned_adam_kernel(const Params *plist) {
    const Params &p = plist[0];
    if (!training(p)) return;
    net::NetBuffers n = p.net;
    net::adam_body(n, p.net_partials, kGradGroups, p.train_batch, p.train_rate);
}

}  // namespace rt

void launch_learned_infer(const rt::Params *p, uint32_t band, hipStream_t s) {
    uint32_t blocks = (uint32_t)std::min<uint64_t>(((uint64_t)band + 63) / 64, 16384);
    hipLaunchKernelGGL(rt::learned_infer_kernel, dim3(blocks ? blocks : 1), dim3(64), 0, s, p);
}

void launch_learned_infer_first(const rt::Params *p, uint32_t band, hipStream_t s) {
    uint32_t blocks = (uint32_t)std::min<uint64_t>(((uint64_t)band + 63) / 64, 16384);
    hipLaunchKernelGGL(rt::learned_infer_first_kernel, dim3(blocks ? blocks : 1), dim3(64), 0, s, p);
}

void launch_learned_train(const rt::Params *p, bool spectral, uint32_t steps, hipStream_t s) {
    for (uint32_t step = 0; step < steps; step++) {
        if (spectral) {
            hipLaunchKernelGGL(rt::learned_weights_kernel<true>, dim3(net::kWeightBlocks), dim3(256), 0, s, p, step);
            hipLaunchKernelGGL(rt::learned_grad_kernel<true>, dim3(kGradGroups), dim3(net::kGradThreads), 0, s, p, step);
        } else {
            hipLaunchKernelGGL(rt::learned_weights_kernel<false>, dim3(net::kWeightBlocks), dim3(256), 0, s, p, step);
            hipLaunchKernelGGL(rt::learned_grad_kernel<false>, dim3(kGradGroups), dim3(net::kGradThreads), 0, s, p, step);
        }
        hipLaunchKernelGGL(rt::learned_adam_kernel, dim3(net::kAdamBlocks), dim3(256), 0, s, p);
    }
}

}  // namespace lr
