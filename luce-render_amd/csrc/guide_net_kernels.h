// Kernels of the learned guide (guide_net.h): inference of 16 points per
// wave, and training steps (weight sum, gradients by backpropagation on the
// matrix cores, Adam). Each is templated on where its points or samples come
// from: host-provided arrays (the library's test entry points) or the
// integrator's path records.
#pragma once

#include "guide_net.h"

namespace lr {
namespace net {

// A training sample: a surface point, its normal, the direction a path left
// it by and the weight value / pdf of that direction (the incident radiance
// estimate over the sampling density).
struct Sample {
    float p[3], n[3], d[3], weight;
    bool valid;
};

constexpr float kMaxSampleScale = 1024.0f;  // per-sample weight over the batch mean, at most
constexpr float kAdamBeta1 = 0.9f, kAdamBeta2 = 0.99f, kAdamEps = 1e-8f;

// Loss of one sample, -c log q(d); the 32 outputs in `io` are replaced by
// the loss's gradient by them (the 12 unused get 0).
__device__ inline float sample_loss(float *io, const float *d, float c) {
    Mixture mx = mixture(io);
    float dens[kLobes], cosk[kLobes], len[kLobes], lk[kLobes];
    float pdf = 0.0f;
    for (uint32_t k = 0; k < kLobes; k++) {
        cosk[k] = mx.mu[k][0] * d[0] + mx.mu[k][1] * d[1] + mx.mu[k][2] * d[2];
        dens[k] = mx.w[k] * vmf(mx.kappa[k], cosk[k]);
        pdf += dens[k];
        float x = io[5 * k], y = io[5 * k + 1], z = io[5 * k + 2];
        len[k] = sqrtf(x * x + y * y + z * z);
        lk[k] = io[5 * k + 3];
    }
    for (uint32_t i = 0; i < kOut; i++) io[i] = 0.0f;
    if (!(pdf > 1e-30f) || c == 0.0f) return 0.0f;
    for (uint32_t k = 0; k < kLobes; k++) {
        float resp = dens[k] / pdf;
        float kappa = mx.kappa[k];
        io[5 * k + 4] = -c * (resp - mx.w[k]);
        if (lk[k] > kLogKappaLo && lk[k] < kLogKappaHi) {
            float t = 2.0f * kappa / expm1f(2.0f * kappa);
            io[5 * k + 3] = -c * resp * (1.0f - t + kappa * (cosk[k] - 1.0f));
        }
        if (len[k] >= 1e-6f) {
            float s = -c * resp * kappa / len[k];
            for (uint32_t a = 0; a < 3u; a++) io[5 * k + a] = s * (d[a] - cosk[k] * mx.mu[k][a]);
        }
    }
    return -c * logf(pdf);
}

// Workgroup-wide (256 threads): adds the valid samples' weights into
// stats[0] and counts the Adam step in stats[1].
template <class Source> __device__ __forceinline__ void weights_body(NetBuffers &net, Source &src, uint32_t n) {
    __shared__ float part[256];
    float sum = 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        Sample s = src.sample(i);
        if (s.valid) sum += s.weight;
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t w = blockDim.x / 2u; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(&net.stats[0], part[0]);
        if (blockIdx.x == 0) atomicAdd(&net.stats[1], 1.0f);
    }
}

template <class Source>
__global__ __launch_bounds__(256) void weights_kernel(NetBuffers net, Source src, uint32_t n) {
    weights_body(net, src, n);
}

// One wave a workgroup. Each wave takes groups of 16 samples (stride: the
// grid) and adds its gradient sums into row blockIdx.x % rows of `partials`
// (kParams + 1 floats: the last the loss), which must start at zero. The
// adds are atomics whose results nobody waits for, so a wave never stalls on
// them. Pre-activations wait in LDS for the backward pass, keeping the
// registers to one layer's tiles.
template <class Source>
__device__ __forceinline__ void grad_body(NetBuffers &net, Source &src, uint32_t n, float *partials, uint32_t rows) {
    __shared__ float xs[16][kIn];          // the encodings
    __shared__ float zs[3][16][kHidden];   // hidden pre-activations
    __shared__ float ds[16][kHidden];      // dZ of the layer being differentiated
    __shared__ float os[16][kOut];         // outputs, then their gradients
    const uint32_t lane = threadIdx.x, j = lane % 16u, g = lane / 16u;
    float *part = partials + (size_t)(blockIdx.x % rows) * (kParams + 1u);
    float mean = net.stats[0] / (float)(n > 0 ? n : 1);
    float loss = 0.0f;
    for (uint32_t base = blockIdx.x * 16u; base < n; base += gridDim.x * 16u) {
        Sample s = src.sample(base + j < n ? base + j : n - 1u);
        bool valid = base + j < n && s.valid && mean > 0.0f;
        float c = valid ? fminf(s.weight / mean, kMaxSampleScale) : 0.0f;
        half4_t h[4];
        encode(net.bounds, s.p[0], s.p[1], s.p[2], s.n[0], s.n[1], s.n[2], g, h);
#pragma unroll
        for (uint32_t m = 0; m < 4u; m++)
#pragma unroll
            for (uint32_t r = 0; r < 4u; r++) xs[j][16u * m + 4u * g + r] = (float)h[m][r];
        // forward, layer by layer
#pragma unroll 1
        for (uint32_t l = 0; l < 3u; l++) {
            float4_t z[4];
            layer_forward(net.fwd, net.params, l, h, lane, z);
#pragma unroll
            for (uint32_t m = 0; m < 4u; m++) {
                float4_t a = z[m];
#pragma unroll
                for (uint32_t r = 0; r < 4u; r++) {
                    zs[l][j][16u * m + 4u * g + r] = a[r];
                    a[r] = fmaxf(a[r], 0.0f);
                }
                h[m] = to_half(a);
            }
        }
        {
            float4_t o[4];
            layer_forward(net.fwd, net.params, 3u, h, lane, o);
#pragma unroll
            for (uint32_t m = 0; m < 2u; m++)
#pragma unroll
                for (uint32_t r = 0; r < 4u; r++) os[j][16u * m + 4u * g + r] = o[m][r];
        }
        __syncthreads();
        // one lane per sample forms the loss gradient
        if (g == 0u) loss += sample_loss(os[j], s.d, c);
        __syncthreads();
        // dZ of the output layer, as tiles (D layout)
        float4_t dz[4];
#pragma unroll
        for (uint32_t m = 0; m < 2u; m++)
#pragma unroll
            for (uint32_t r = 0; r < 4u; r++) dz[m][r] = os[j][16u * m + 4u * g + r];
        dz[2] = float4_t{0.0f, 0.0f, 0.0f, 0.0f};
        dz[3] = dz[2];
#pragma unroll 1
        for (int l = 3; l >= 0; l--) {
            const uint32_t outs = layer_out((uint32_t)l), steps = outs / 16u;
#pragma unroll
            for (uint32_t m = 0; m < 4u; m++)
                if (m < steps)
#pragma unroll
                    for (uint32_t r = 0; r < 4u; r++) ds[j][16u * m + 4u * g + r] = dz[m][r];
            __syncthreads();
            // dW = H^T dZ: rows (inputs) by tile m, columns (outputs) by tile q,
            // the 16 samples as the product's inner dimension; H the layer's
            // input (the encoding, or the ReLU of the previous pre-activation)
#pragma unroll 1
            for (uint32_t m = 0; m < 4u; m++) {
                half4_t a;
#pragma unroll
                for (uint32_t r = 0; r < 4u; r++)
                    a[r] = (_Float16)(l == 0 ? xs[4u * g + r][16u * m + j] : fmaxf(zs[l - 1][4u * g + r][16u * m + j], 0.0f));
                for (uint32_t q = 0; q < steps; q++) {
                    half4_t b;
#pragma unroll
                    for (uint32_t r = 0; r < 4u; r++) b[r] = (_Float16)ds[4u * g + r][16u * q + j];
                    float4_t w = mfma(a, b, float4_t{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
                    for (uint32_t r = 0; r < 4u; r++)
                        atomicAdd(part + weight_offset((uint32_t)l) + (16u * m + 4u * g + r) * outs + 16u * q + j, w[r]);
                }
            }
            // db = sum of dZ over the samples
            if (lane < outs) {
                float sb = 0.0f;
                for (uint32_t k = 0; k < 16u; k++) sb += ds[k][lane];
                atomicAdd(&part[bias_offset((uint32_t)l) + lane], sb);
            }
            if (l > 0) {
                // dH^T = W dZ^T (K = 32 blocks of outputs), then through the
                // ReLU of layer l - 1
                const uint32_t blocks = outs / 32u;
                const half8_t b0 = pair(to_half(dz[0]), to_half(dz[1])), b1 = pair(to_half(dz[2]), to_half(dz[3]));
#pragma unroll
                for (uint32_t m = 0; m < 4u; m++) {
                    float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
                    acc = mfma32(load_block(net.bwd, block_base((uint32_t)l) + m * blocks, lane), b0, acc);
                    if (blocks > 1u)
                        acc = mfma32(load_block(net.bwd, block_base((uint32_t)l) + m * blocks + 1u, lane), b1, acc);
#pragma unroll
                    for (uint32_t r = 0; r < 4u; r++) dz[m][r] = zs[l - 1][j][16u * m + 4u * g + r] > 0.0f ? acc[r] : 0.0f;
                }
            }
            __syncthreads();
        }
    }
    // the wave's loss
    for (uint32_t off = 32u; off > 0; off >>= 1) loss += __shfl_xor(loss, (int)off);
    if (lane == 0) atomicAdd(&part[kParams], loss);
}

template <class Source>
__global__ __launch_bounds__(64) void grad_kernel(NetBuffers net, Source src, uint32_t n, float *partials,
                                                  uint32_t rows) {
    grad_body(net, src, n, partials, rows);
}

// Writes parameter i (f32) and its FP16 fragments.
__device__ __forceinline__ void store_param(NetBuffers &net, uint32_t i, float w) {
    net.params[i] = w;
    if (i >= kBias) return;
    uint32_t l = i < kW3 ? i / (kIn * kHidden) : 3u;
    uint32_t local = i - weight_offset(l), outs = layer_out(l);
    uint32_t in = local / outs, out = local % outs;
    net.fwd[fwd_index(l, in, out)] = (_Float16)w;
    net.bwd[bwd_index(l, in, out)] = (_Float16)w;
}

// One thread a parameter (kParams + 1 threads when clearing): the gradient
// is the sum of `waves` partial rows
// over the samples; an Adam step at rate `rate` (0: only refresh the FP16
// fragments from the master copy). Clears the batch's weight sum.
// With `clear`, the partial rows read are zeroed for the next batch.
__device__ __forceinline__ void adam_body(NetBuffers &net, float *partials, uint32_t waves, uint32_t samples,
                                          float rate, bool clear) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (clear && i == kParams)
        for (uint32_t k = 0; k < waves; k++) partials[(size_t)k * (kParams + 1u) + kParams] = 0.0f;
    if (i < kParams) {
        float w = net.params[i];
        if (rate > 0.0f) {
            float grad = 0.0f;
            for (uint32_t k = 0; k < waves; k++) {
                float *at = partials + (size_t)k * (kParams + 1u) + i;
                grad += *at;
                if (clear) *at = 0.0f;
            }
            grad /= (float)(samples > 0 ? samples : 1);
            float t = net.stats[1];
            float m = kAdamBeta1 * net.m[i] + (1.0f - kAdamBeta1) * grad;
            float v = kAdamBeta2 * net.v[i] + (1.0f - kAdamBeta2) * grad * grad;
            net.m[i] = m;
            net.v[i] = v;
            float mh = m / (1.0f - powf(kAdamBeta1, t)), vh = v / (1.0f - powf(kAdamBeta2, t));
            w -= rate * mh / (sqrtf(vh) + kAdamEps);
        }
        store_param(net, i, w);
    }
    if (i == 0) net.stats[0] = 0.0f;
}


// Mixtures of points, 64 a wave at a time: each lane gathers one point and
// encodes it into its row of LDS, the matrix cores take the rows 16 at a
// time (four forward passes), each pass's outputs replace its rows' encodings
// (a row holds 64 halves, then 32 floats: one wave owns the buffer and its
// LDS operations complete in order), and each lane forms its own point's
// mixture. `dst.point(i)` gives point i's position and normal,
// `dst.store(i, mixture)` takes the result.
constexpr uint32_t kRowWords = 36;  // 32 words a row and 4 of padding: rows 16-byte aligned, 16 rows read at one column 2 to a bank
template <class Points> __device__ __forceinline__ void infer_body(const NetBuffers &net, Points &pts, uint32_t n) {
    __shared__ uint32_t rows[64 * kRowWords];
    const uint32_t lane = threadIdx.x, j = lane % 16u, g = lane / 16u;
    _Float16 *own = reinterpret_cast<_Float16 *>(rows + lane * kRowWords);
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        float p[3], nn[3];
        pts.point(base + lane < n ? base + lane : n - 1u, p, nn);
        encode_row(net.bounds, p[0], p[1], p[2], nn[0], nn[1], nn[2], own);
        __syncthreads();
#pragma unroll 1
        for (uint32_t q = 0; q < 4u; q++) {
            const _Float16 *row = reinterpret_cast<const _Float16 *>(rows + (16u * q + j) * kRowWords);
            half4_t x[4];
#pragma unroll
            for (uint32_t kk = 0; kk < 4u; kk++) x[kk] = *reinterpret_cast<const half4_t *>(row + 16u * kk + 4u * g);
            Act z[3];
            float4_t o[2];
            forward(net.fwd, net.params, x, lane, z, o);
            float *out = reinterpret_cast<float *>(rows + (16u * q + j) * kRowWords);
#pragma unroll
            for (uint32_t m = 0; m < 2u; m++) *reinterpret_cast<float4_t *>(out + 16u * m + 4u * g) = o[m];
        }
        __syncthreads();
        if (base + lane < n) pts.store(base + lane, mixture(reinterpret_cast<const float *>(rows + lane * kRowWords)));
        __syncthreads();
    }
}

template <class Points>
__global__ __launch_bounds__(64) void infer_kernel(NetBuffers net, Points pts, uint32_t n) {
    infer_body(net, pts, n);
}

}  // namespace net
}  // namespace lr
