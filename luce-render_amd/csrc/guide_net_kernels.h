// Kernels of the learned guide (guide_net.h): inference of 64 points per
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
// Inference reads an exponential moving average of the trained parameters
// (weight kEmaBeta on the average so far, bias-corrected like Adam's
// moments): one step's noise does not set the guide.
constexpr float kEmaBeta = 0.8f;

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

// Workgroup-wide (256 threads, at most kWeightBlocks workgroups): stores the
// sum of its valid samples' weights in its stats word (the block count at
// stats[2]) and counts the Adam step in stats[1].
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
        net.stats[kWeightSums + blockIdx.x] = part[0];
        if (blockIdx.x == 0) {
            net.stats[2] = (float)gridDim.x;
            net.stats[1] += 1.0f;
        }
    }
}

// The batch's weight sum: the weight-sum kernel's block sums in block order.
__device__ __forceinline__ float batch_weight(const NetBuffers &net) {
    const uint32_t blocks = (uint32_t)net.stats[2];
    float total = 0.0f;
    for (uint32_t b = 0; b < blocks; b++) total += net.stats[kWeightSums + b];
    return total;
}

template <class Source>
__global__ __launch_bounds__(256) void weights_kernel(NetBuffers net, Source src, uint32_t n) {
    weights_body(net, src, n);
}

// One workgroup of four waves a pair of 16-sample tiles (32 samples;
// stride: the grid), its gradient sums written into its own row, blockIdx.x,
// of `partials` (kParams + 1 floats: the last the loss): plain stores for its
// first pair, adds for later ones, so rows need no clearing and no workgroup
// waits on another (the grid must not exceed the rows). The waves share the
// pair through LDS and split every layer by 16-row tiles: wave w forms
// output tile w of a forward product, and input tile w of dW = H^T dZ (the
// 32 samples as the inner dimension of one K = 32 product) and of
// dH^T = W dZ^T; a barrier ends each layer. Fragments are built as the
// single-tile products take them (two D tiles side by side as a K = 32 B
// operand), so the sums are the same; encodings and hidden pre-activations
// are kept as the halves the products take.
constexpr uint32_t kGradThreads = 256u;
template <class Source>
__device__ __forceinline__ void grad_body(NetBuffers &net, Source &src, uint32_t n, float *partials) {
    constexpr uint32_t kPair = 32u;
    __shared__ __attribute__((aligned(16))) _Float16 xs[kPair][kIn];
    __shared__ __attribute__((aligned(16))) _Float16 zs[3][kPair][kHidden];
    __shared__ __attribute__((aligned(16))) float ds[2][kPair][kHidden];  // dZ of the layer differentiated, then the one below
    __shared__ float os[kPair][kOut];        // outputs, then their gradients
    __shared__ float dir[kPair][3];
    __shared__ float cw[kPair];
    const uint32_t wave = threadIdx.x / 64u, lane = threadIdx.x % 64u, j = lane % 16u, g = lane / 16u;
    float *part = partials + (size_t)blockIdx.x * (kParams + 1u);
    const float mean = batch_weight(net) / (float)(n > 0 ? n : 1);
    // layer l's input k of pair row `row`: the encoding, or the ReLU of the
    // previous pre-activation
    auto input = [&](uint32_t l, uint32_t row, uint32_t k) -> _Float16 {
        if (l == 0u) return xs[row][k];
        _Float16 z = zs[l - 1u][row][k];
        return z > (_Float16)0.0f ? z : (_Float16)0.0f;
    };
    float loss = 0.0f;
    bool first = true;
    for (uint32_t base = blockIdx.x * kPair; base < n; base += gridDim.x * kPair, first = false) {
        // waves 0 and 1 gather and encode tile `wave` (lane (j, g): sample j)
        if (wave < 2u) {
            const uint32_t row = 16u * wave + j, i = base + row;
            Sample s = src.sample(i < n ? i : n - 1u);
            bool valid = i < n && s.valid && mean > 0.0f;
            half4_t h[4];
            encode(net.bounds, s.p[0], s.p[1], s.p[2], s.n[0], s.n[1], s.n[2], g, h);
#pragma unroll
            for (uint32_t m = 0; m < 4u; m++)
#pragma unroll
                for (uint32_t r = 0; r < 4u; r++) xs[row][16u * m + 4u * g + r] = h[m][r];
            if (g == 0u) {
                dir[row][0] = s.d[0], dir[row][1] = s.d[1], dir[row][2] = s.d[2];
                cw[row] = valid ? fminf(s.weight / mean, kMaxSampleScale) : 0.0f;
            }
        }
        __syncthreads();
        // forward: wave w forms output tile w of each layer, for both tiles
#pragma unroll 1
        for (uint32_t l = 0; l < 4u; l++) {
            if (wave < layer_out(l) / 16u) {
                const uint32_t m = wave;
                const half8_t a0 = load_block(net.fwd, block_base(l) + m * 2u, lane);
                const half8_t a1 = load_block(net.fwd, block_base(l) + m * 2u + 1u, lane);
                const float *b = net.params + bias_offset(l) + m * 16u + 4u * g;
#pragma unroll
                for (uint32_t t = 0; t < 2u; t++) {
                    const uint32_t row = 16u * t + j;
                    // the four inputs 16c + 4g .. +3 of each 16-input tile c, one 8-byte read each
                    const _Float16 *in = l == 0u ? &xs[row][0] : &zs[l - 1u][row][0];
                    half4_t c[4];
#pragma unroll
                    for (uint32_t k = 0; k < 4u; k++) {
                        c[k] = *reinterpret_cast<const half4_t *>(in + 16u * k + 4u * g);
                        if (l != 0u)
#pragma unroll
                            for (uint32_t r = 0; r < 4u; r++) c[k][r] = c[k][r] > (_Float16)0.0f ? c[k][r] : (_Float16)0.0f;
                    }
                    const half8_t b0 = pair(c[0], c[1]), b1 = pair(c[2], c[3]);
                    float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
                    acc = mfma32(a0, b0, acc);
                    acc = mfma32(a1, b1, acc);
#pragma unroll
                    for (uint32_t r = 0; r < 4u; r++) {
                        float z = acc[r] + b[r];
                        if (l < 3u) zs[l][row][16u * m + 4u * g + r] = (_Float16)z;
                        else os[row][16u * m + 4u * g + r] = z;
                    }
                }
            }
            __syncthreads();
        }
        // wave 0: one lane per sample forms the loss gradient in its row
        if (wave == 0u && g < 2u) {
            const uint32_t row = 16u * g + j;
            float d[3] = {dir[row][0], dir[row][1], dir[row][2]};
            loss += sample_loss(os[row], d, cw[row]);
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < kPair * kOut; k += kGradThreads) ds[0][k / kOut][k % kOut] = os[k / kOut][k % kOut];
        __syncthreads();
        uint32_t cur = 0u;
#pragma unroll 1
        for (int l = 3; l >= 0; l--) {
            const uint32_t outs = layer_out((uint32_t)l), steps = outs / 16u, m = wave;
            float(*dz)[kHidden] = ds[cur];
            {
                half8_t a;
#pragma unroll
                for (uint32_t e = 0; e < 8u; e++) a[e] = input((uint32_t)l, 8u * g + e, 16u * m + j);
                for (uint32_t q = 0; q < steps; q++) {
                    half8_t b;
#pragma unroll
                    for (uint32_t e = 0; e < 8u; e++) b[e] = (_Float16)dz[8u * g + e][16u * q + j];
                    float4_t w = mfma32(a, b, float4_t{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
                    for (uint32_t r = 0; r < 4u; r++) {
                        float *at = part + weight_offset((uint32_t)l) + (16u * m + 4u * g + r) * outs + 16u * q + j;
                        *at = first ? w[r] : *at + w[r];
                    }
                }
            }
            if (wave == 0u && lane < outs) {
                float sb = 0.0f;
                for (uint32_t k = 0; k < kPair; k++) sb += dz[k][lane];
                float *at = part + bias_offset((uint32_t)l) + lane;
                *at = first ? sb : *at + sb;
            }
            if (l > 0) {
                const uint32_t blocks = outs / 32u;
                const half8_t a0 = load_block(net.bwd, block_base((uint32_t)l) + m * blocks, lane);
                const half8_t a1 = blocks > 1u ? load_block(net.bwd, block_base((uint32_t)l) + m * blocks + 1u, lane) : a0;
#pragma unroll
                for (uint32_t t = 0; t < 2u; t++) {
                    const uint32_t row = 16u * t + j;
                    // dZ's outputs 16c + 4g .. +3, one 16-byte read each
                    half4_t c[4];
#pragma unroll
                    for (uint32_t k = 0; k < 4u; k++) {
                        float4_t v = k < 2u || blocks > 1u ? *reinterpret_cast<const float4_t *>(&dz[row][16u * k + 4u * g])
                                                           : float4_t{0.0f, 0.0f, 0.0f, 0.0f};
                        c[k] = to_half(v);
                    }
                    const half8_t b0 = pair(c[0], c[1]), b1 = pair(c[2], c[3]);
                    float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
                    acc = mfma32(a0, b0, acc);
                    if (blocks > 1u) acc = mfma32(a1, b1, acc);
#pragma unroll
                    for (uint32_t r = 0; r < 4u; r++)
                        ds[cur ^ 1u][row][16u * m + 4u * g + r] =
                            zs[l - 1][row][16u * m + 4u * g + r] > (_Float16)0.0f ? acc[r] : 0.0f;
                }
            }
            __syncthreads();
            cur ^= 1u;
        }
    }
    // the workgroup's loss: wave 0's lanes
    if (wave == 0u) {
        for (uint32_t off = 32u; off > 0; off >>= 1) loss += __shfl_xor(loss, (int)off);
        if (lane == 0 && blockIdx.x * kPair < n) part[kParams] = loss;
    }
}

template <class Source>
__global__ __launch_bounds__(kGradThreads) void grad_kernel(NetBuffers net, Source src, uint32_t n, float *partials) {
    grad_body(net, src, n, partials);
}

// Writes parameter i (f32) and its FP16 fragments: `w` the trained value,
// `used` the one inference uses.
__device__ __forceinline__ void store_param(NetBuffers &net, uint32_t i, float w, float used) {
    net.params[i] = w;
    net.infer[i] = used;
    if (i >= kBias) return;
    uint32_t l = i < kW3 ? i / (kIn * kHidden) : 3u;
    uint32_t local = i - weight_offset(l), outs = layer_out(l);
    uint32_t in = local / outs, out = local % outs;
    net.fwd[fwd_index(l, in, out)] = (_Float16)w;
    net.bwd[bwd_index(l, in, out)] = (_Float16)w;
    net.infer_fwd[fwd_index(l, in, out)] = (_Float16)used;
}

// Workgroups of 256 threads, kAdamParams parameters each (kAdamBlocks of
// them): kAdamParts threads a parameter sum every kAdamParts-th of the first
// `waves` partial rows
// (those the gradient kernel wrote), then one takes the gradient over the
// samples and an Adam step at rate `rate`, then the moving average and the
// inference copy; rate 0 only refreshes the fragments from the master copy,
// which inference then uses as it is.
constexpr uint32_t kAdamParts = 8u, kAdamParams = 256u / kAdamParts;
constexpr uint32_t kAdamBlocks = (kParams + kAdamParams - 1u) / kAdamParams;
__device__ __forceinline__ void adam_body(NetBuffers &net, const float *partials, uint32_t waves, uint32_t samples,
                                          float rate) {
    __shared__ float sums[kAdamParts][kAdamParams];
    const uint32_t local = threadIdx.x % kAdamParams, part = threadIdx.x / kAdamParams;
    const uint32_t i = blockIdx.x * kAdamParams + local;
    float grad = 0.0f;
    if (i < kParams && rate > 0.0f) {
#pragma unroll 4
        for (uint32_t k = part; k < waves; k += kAdamParts) grad += partials[(size_t)k * (kParams + 1u) + i];
    }
    sums[part][local] = grad;
    __syncthreads();
    if (part == 0u && i < kParams) {
        float w = net.params[i];
        float t = net.stats[1];
        float used = w;
        if (rate > 0.0f) {
            grad = 0.0f;
#pragma unroll
            for (uint32_t k = 0; k < kAdamParts; k++) grad += sums[k][local];
            grad /= (float)(samples > 0 ? samples : 1);
            float m = kAdamBeta1 * net.m[i] + (1.0f - kAdamBeta1) * grad;
            float v = kAdamBeta2 * net.v[i] + (1.0f - kAdamBeta2) * grad * grad;
            net.m[i] = m;
            net.v[i] = v;
            float mh = m / (1.0f - powf(kAdamBeta1, t)), vh = v / (1.0f - powf(kAdamBeta2, t));
            w -= rate * mh / (sqrtf(vh) + kAdamEps);
            float e = kEmaBeta * net.ema[i] + (1.0f - kEmaBeta) * w;
            net.ema[i] = e;
            used = e / (1.0f - powf(kEmaBeta, t));
        } else {
            // the average restarts from w: its corrected value is w now
            net.ema[i] = w * (1.0f - powf(kEmaBeta, t));
        }
        store_param(net, i, w, used);
    }
}

// Mixtures of points, 64 a wave at a time: each lane gathers one point and
// encodes it into its row of LDS, the rows are read as four 16-point tiles
// that go through the layers together, the outputs replace the rows'
// encodings (a row holds 64 halves, then 32 floats: one wave owns the buffer
// and its LDS operations complete in order), and each lane forms its own
// point's mixture from the moving-average parameters. `pts.point(i)` gives
// point i's position and normal, `pts.store(i, mixture)` takes the result.
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
        // The four 16-point tiles' input fragments, then the layers one at a
        // time over all four, each weight block loaded once for them
        // (layer_forward's products and ReLU, tile by tile).
        half4_t h[4][4];
#pragma unroll
        for (uint32_t q = 0; q < 4u; q++) {
            const _Float16 *row = reinterpret_cast<const _Float16 *>(rows + (16u * q + j) * kRowWords);
#pragma unroll
            for (uint32_t kk = 0; kk < 4u; kk++) h[q][kk] = *reinterpret_cast<const half4_t *>(row + 16u * kk + 4u * g);
        }
#pragma unroll 1
        for (uint32_t l = 0; l < 3u; l++) {
            half4_t next[4][4];
#pragma unroll
            for (uint32_t m = 0; m < 4u; m++) {
                const half8_t a0 = load_block(net.infer_fwd, block_base(l) + m * 2u, lane);
                const half8_t a1 = load_block(net.infer_fwd, block_base(l) + m * 2u + 1u, lane);
                const float *b = net.infer + bias_offset(l) + m * 16u + 4u * g;
#pragma unroll
                for (uint32_t q = 0; q < 4u; q++) {
                    float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
                    acc = mfma32(a0, pair(h[q][0], h[q][1]), acc);
                    acc = mfma32(a1, pair(h[q][2], h[q][3]), acc);
#pragma unroll
                    for (int r = 0; r < 4; r++) acc[r] = fmaxf(acc[r] + b[r], 0.0f);
                    next[q][m] = to_half(acc);
                }
            }
#pragma unroll
            for (uint32_t q = 0; q < 4u; q++)
#pragma unroll
                for (uint32_t m = 0; m < 4u; m++) h[q][m] = next[q][m];
        }
        // the output layer's two tiles, into the rows (the inputs are all read)
#pragma unroll
        for (uint32_t m = 0; m < 2u; m++) {
            const half8_t a0 = load_block(net.infer_fwd, block_base(3u) + m * 2u, lane);
            const half8_t a1 = load_block(net.infer_fwd, block_base(3u) + m * 2u + 1u, lane);
            const float *b = net.infer + bias_offset(3u) + m * 16u + 4u * g;
#pragma unroll
            for (uint32_t q = 0; q < 4u; q++) {
                float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
                acc = mfma32(a0, pair(h[q][0], h[q][1]), acc);
                acc = mfma32(a1, pair(h[q][2], h[q][3]), acc);
#pragma unroll
                for (int r = 0; r < 4; r++) acc[r] += b[r];
                float *out = reinterpret_cast<float *>(rows + (16u * q + j) * kRowWords);
                *reinterpret_cast<float4_t *>(out + 16u * m + 4u * g) = acc;
            }
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
