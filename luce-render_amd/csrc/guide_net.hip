// The learned guide's device state and its C ABI (test and standalone use):
// create with a scene's bounds, read and write the master parameters, the
// gradient of a batch, Adam steps and mixtures of points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "guide_net_kernels.h"
#include "lr_internal.h"
#include "lr_scene.h"
#include "guide_net_host.h"

using namespace lr;
using namespace lr::net;

namespace {

// Samples as the ABI takes them: 10 floats each (position, normal,
// direction, weight).
struct ArraySamples {
    const float *a;
    __device__ __forceinline__ Sample sample(uint32_t i) const {
        const float *x = a + (size_t)i * 10u;
        Sample s;
        for (int k = 0; k < 3; k++) s.p[k] = x[k], s.n[k] = x[3 + k], s.d[k] = x[6 + k];
        s.weight = x[9];
        s.valid = x[9] > 0.0f;
        return s;
    }
};

// Points as the ABI takes them (position, normal: 6 floats); mixtures out as
// 20 floats (per lobe direction and concentration, then the 4 weights).
struct ArrayPoints {
    const float *a;
    float *out;
    __device__ __forceinline__ void point(uint32_t i, float *p, float *n) const {
        const float *x = a + (size_t)i * 6u;
        for (int k = 0; k < 3; k++) p[k] = x[k], n[k] = x[3 + k];
    }
    __device__ __forceinline__ void store(uint32_t i, const Mixture &m) const {
        float *o = out + (size_t)i * 20u;
        for (uint32_t k = 0; k < kLobes; k++) {
            o[4 * k] = m.mu[k][0], o[4 * k + 1] = m.mu[k][1], o[4 * k + 2] = m.mu[k][2], o[4 * k + 3] = m.kappa[k];
            o[16 + k] = m.w[k];
        }
    }
};

__global__ __launch_bounds__(256) void adam_kernel(NetBuffers net, float *partials, uint32_t waves, uint32_t samples,
                                                   float rate) {
    adam_body(net, partials, waves, samples, rate);
}

__global__ void zero_kernel(float *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0f;
}

}  // namespace

namespace lr {

GuideNet::~GuideNet() {
    for (void *p : buffers) (void)hipFree(p);
}

int GuideNet::create(lr_device *dev, const float lo[3], const float hi[3], uint32_t seed, uint32_t max_waves) {
    device = dev;
    waves = max_waves;
    auto alloc = [&](size_t bytes) -> void * {
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        buffers.push_back(p);
        return p;
    };
    b.params = (float *)alloc(kParams * 4);
    b.m = (float *)alloc(kParams * 4);
    b.v = (float *)alloc(kParams * 4);
    b.fwd = (_Float16 *)alloc(kFragmentHalves * 2);
    b.bwd = (_Float16 *)alloc(kFragmentHalves * 2);
    b.ema = (float *)alloc(kParams * 4);
    b.infer = (float *)alloc(kParams * 4);
    b.infer_fwd = (_Float16 *)alloc(kFragmentHalves * 2);
    b.bounds = (float *)alloc(64);
    b.stats = (float *)alloc(kStatsWords * 4);
    partials = (float *)alloc((size_t)waves * (kParams + 1) * 4);
    if (!b.params || !b.m || !b.v || !b.fwd || !b.bwd || !b.ema || !b.infer || !b.infer_fwd || !b.bounds || !b.stats ||
        !partials)
        return fail(LR_DEVICE_ERROR, "hipMalloc(learned guide)");
    // He-uniform initial weights from a fixed integer hash; zero biases.
    std::vector<float> init(kParams, 0.0f);
    uint64_t state = 0x9E3779B97F4A7C15ull ^ ((uint64_t)seed << 17);
    for (uint32_t l = 0; l < kLayers; l++) {
        float limit = std::sqrt(6.0f / (float)kIn);
        uint32_t count = kIn * layer_out(l);
        for (uint32_t i = 0; i < count; i++) {
            state = state * 6364136223846793005ull + 1442695040888963407ull;
            float u = (float)((state >> 40) & 0xFFFFFF) / 16777216.0f;
            init[weight_offset(l) + i] = (2.0f * u - 1.0f) * limit;
        }
    }
    float bounds[16] = {};
    for (int k = 0; k < 3; k++) {
        float extent = hi[k] - lo[k];
        bounds[k] = lo[k];
        bounds[3 + k] = extent > 0.0f ? 1.0f / extent : 1.0f;
    }
    float stats[kStatsWords] = {};
    hipStream_t s = dev->stream;
    hipError_t e = hipMemcpyAsync(b.bounds, bounds, 64, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(b.stats, stats, sizeof(stats), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "learned guide set-up");
    hipLaunchKernelGGL(zero_kernel, dim3(256), dim3(256), 0, s, partials, (size_t)waves * (kParams + 1));
    return set_params(init.data());
}

int GuideNet::set_params(const float *params) {
    hipStream_t s = device->stream;
    hipError_t e = hipMemcpyAsync(b.params, params, kParams * 4, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "learned guide parameters");
    hipLaunchKernelGGL(zero_kernel, dim3(64), dim3(256), 0, s, b.m, (size_t)kParams);
    hipLaunchKernelGGL(zero_kernel, dim3(64), dim3(256), 0, s, b.v, (size_t)kParams);
    refresh(s);
    return 0;
}

void GuideNet::refresh(hipStream_t s) {
    hipLaunchKernelGGL(adam_kernel, dim3(kAdamBlocks), dim3(256), 0, s, b, partials, 0u, 0u, 0.0f);
}

}  // namespace lr

extern "C" {

int lr_guide_net_create(lr_device *device, const float lo[3], const float hi[3], uint32_t seed, lr_guide_net **out) {
    if (!device || !lo || !hi || !out) return fail(LR_INVALID, "lr_guide_net_create needs its arguments");
    *out = nullptr;
    (void)hipSetDevice(device->ordinal);
    std::unique_ptr<lr_guide_net> n(new lr_guide_net());
    int rc = n->net.create(device, lo, hi, seed, 256);
    if (rc) return rc;
    hipError_t e = hipStreamSynchronize(device->stream);
    if (e != hipSuccess) return hip_fail(e, "learned guide set-up");
    *out = n.release();
    return 0;
}

void lr_guide_net_destroy(lr_guide_net *net) {
    if (!net) return;
    (void)hipStreamSynchronize(net->net.device->stream);
    delete net;
}

size_t lr_guide_net_param_count(void) { return kParams; }

int lr_guide_net_set_params(lr_guide_net *net, const float *params, size_t count) {
    if (!net || !params) return fail(LR_INVALID, "lr_guide_net_set_params needs a network and parameters");
    if (count != kParams) return fail(LR_INVALID, "the learned guide has lr_guide_net_param_count() parameters");
    int rc = net->net.set_params(params);
    if (rc) return rc;
    hipError_t e = hipStreamSynchronize(net->net.device->stream);
    return e == hipSuccess ? 0 : hip_fail(e, "learned guide parameters");
}

int lr_guide_net_get_params(lr_guide_net *net, float *params, size_t count) {
    if (!net || !params) return fail(LR_INVALID, "lr_guide_net_get_params needs a network and an output");
    if (count != kParams) return fail(LR_INVALID, "the learned guide has lr_guide_net_param_count() parameters");
    hipStream_t s = net->net.device->stream;
    hipError_t e = hipMemcpyAsync(params, net->net.b.params, kParams * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? 0 : hip_fail(e, "learned guide parameters");
}

// Gradient (and summed loss) of a batch; with rate > 0 also an Adam step.
static int batch(lr_guide_net *net, const float *samples, size_t n, float rate, float *grad, size_t grad_count,
                 float *loss) {
    if (!net || (n && !samples)) return fail(LR_INVALID, "a learned-guide batch needs a network and samples");
    if (n == 0 || n > (1u << 24)) return fail(LR_INVALID, "a batch holds 1 to 2^24 samples");
    if (grad && grad_count != kParams) return fail(LR_INVALID, "the gradient has lr_guide_net_param_count() entries");
    GuideNet &g = net->net;
    hipStream_t s = g.device->stream;
    void *dev = nullptr;
    hipError_t e = hipMalloc(&dev, n * 40);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(batch)");
    e = hipMemcpyAsync(dev, samples, n * 40, hipMemcpyHostToDevice, s);
    // one workgroup (and partial row) per 32 samples, at most the rows
    // allocated; every workgroup of the grid takes at least one pair of tiles
    uint32_t rows = (uint32_t)std::min<size_t>(g.waves, (n + 31) / 32);
    ArraySamples src{(const float *)dev};
    if (e == hipSuccess) {
        hipLaunchKernelGGL((weights_kernel<ArraySamples>), dim3(kWeightBlocks), dim3(256), 0, s, g.b, src, (uint32_t)n);
        hipLaunchKernelGGL((grad_kernel<ArraySamples>), dim3(rows), dim3(kGradThreads), 0, s, g.b, src, (uint32_t)n, g.partials);
        e = hipGetLastError();
    }
    std::vector<float> part;
    if (e == hipSuccess && (grad || loss)) {
        part.resize((size_t)rows * (kParams + 1));
        e = hipMemcpyAsync(part.data(), g.partials, part.size() * 4, hipMemcpyDeviceToHost, s);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(adam_kernel, dim3(kAdamBlocks), dim3(256), 0, s, g.b, g.partials, rows,
                           (uint32_t)n, rate);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(dev);
    if (e != hipSuccess) return hip_fail(e, "learned guide batch");
    if (grad) {
        for (size_t i = 0; i < kParams; i++) {
            double sum = 0.0;
            for (uint32_t w = 0; w < rows; w++) sum += part[(size_t)w * (kParams + 1) + i];
            grad[i] = (float)(sum / (double)n);
        }
    }
    if (loss) {
        double sum = 0.0;
        for (uint32_t w = 0; w < rows; w++) sum += part[(size_t)w * (kParams + 1) + kParams];
        *loss = (float)(sum / (double)n);
    }
    return 0;
}

int lr_guide_net_gradient(lr_guide_net *net, const float *samples, size_t n, float *grad, size_t count, float *loss) {
    if (!grad) return fail(LR_INVALID, "lr_guide_net_gradient needs an output");
    return batch(net, samples, n, 0.0f, grad, count, loss);
}

int lr_guide_net_step(lr_guide_net *net, const float *samples, size_t n, float rate, float *loss) {
    if (!(rate > 0.0f)) return fail(LR_INVALID, "a learning rate above 0");
    return batch(net, samples, n, rate, nullptr, 0, loss);
}

int lr_guide_net_mixtures(lr_guide_net *net, const float *points, size_t n, float *out) {
    if (!net || (n && (!points || !out))) return fail(LR_INVALID, "lr_guide_net_mixtures needs its arguments");
    if (n == 0) return 0;
    if (n > (1u << 26)) return fail(LR_INVALID, "at most 2^26 points a call");
    GuideNet &g = net->net;
    hipStream_t s = g.device->stream;
    void *dp = nullptr, *dout = nullptr;
    hipError_t e = hipMalloc(&dp, n * 24);
    if (e == hipSuccess) e = hipMalloc(&dout, n * 80);
    if (e == hipSuccess) e = hipMemcpyAsync(dp, points, n * 24, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        uint32_t blocks = (uint32_t)std::min<size_t>((n + 63) / 64, 16384);
        ArrayPoints pts{(const float *)dp, (float *)dout};
        hipLaunchKernelGGL((infer_kernel<ArrayPoints>), dim3(blocks), dim3(64), 0, s, g.b, pts, (uint32_t)n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, n * 80, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (dp) (void)hipFree(dp);
    if (dout) (void)hipFree(dout);
    return e == hipSuccess ? 0 : hip_fail(e, "learned guide mixtures");
}

}  // extern "C"
