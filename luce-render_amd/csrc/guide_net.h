// The learned path guide: a small MLP (64 one-blob inputs -> 3 hidden layers
// of 64 ReLU units -> 32 outputs) that maps a surface point and its normal to
// a mixture of four von Mises-Fisher lobes over incident directions, the same
// distribution family the grid guide (render_guide.h) stores per cell. The
// matrix products run on the CDNA4 matrix cores (v_mfma_f32_16x16x32_f16:
// FP16 operands, FP32 accumulation, K = 32), 16 points per product. No
// reference counterpart: the reference guides with the grid only; this
// network is trained online from the same path records the grid learns
// from, and judged against the grid.
//
// Fragment layouts of v_mfma_f32_16x16x32_f16 (64 lanes; lane = c + 16 g),
// checked on the device:
//   A[i][k] (16x32): lane i + 16 (k / 8), element k % 8
//   B[k][j] (32x16): lane j + 16 (k / 8), element k % 8
//   D[i][j] (16x16): lane j + 16 (i / 4), element i % 4
// A layer computes Z^T = W^T H^T over 16 points (columns): A = W^T blocks,
// B = H^T. A D tile holds rows 4g..4g+3 in lane group g, so two D tiles
// (rows 0..15 and 16..31 of a 32-row block) side by side are a K = 32 B
// fragment whose element e of lane group g is row 4g + e (e < 4) or
// 16 + 4g + e - 4: the A blocks are packed in that order of K, so in
// inference activations never leave their lanes (training's waves share
// them through LDS and build the same operands from it). The weight
// gradient takes 32 samples as the inner dimension of the K = 32 product.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lr {
namespace net {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr uint32_t kIn = 64, kHidden = 64, kOut = 32, kLobes = 4, kLayers = 4;
constexpr uint32_t kPosBins = 16, kNormalBins = 8;
// Parameters (f32 master copy): W0, W1, W2 [64][64], W3 [64][32] (row = input,
// column = output), then b0, b1, b2 [64], b3 [32].
constexpr uint32_t kW3 = 3u * kIn * kHidden;
constexpr uint32_t kBias = kW3 + kHidden * kOut;
constexpr uint32_t kParams = kBias + 3u * kHidden + kOut;  // 14560
// 16x32 FP16 blocks (a half8 a lane): forward A = W^T (16-output tiles x
// 32-input blocks: 8 a hidden layer, 4 the output layer), backward A = W
// (16-input tiles x 32-output blocks: 8, then 4).
constexpr uint32_t kBlocks = 3u * 8u + 4u;  // 28
constexpr size_t kFragmentHalves = (size_t)kBlocks * 64u * 8u;
__host__ __device__ constexpr uint32_t layer_out(uint32_t l) { return l == 3u ? kOut : kHidden; }
__host__ __device__ constexpr uint32_t weight_offset(uint32_t l) { return l * kIn * kHidden; }
__host__ __device__ constexpr uint32_t bias_offset(uint32_t l) { return kBias + l * kHidden; }
__host__ __device__ constexpr uint32_t block_base(uint32_t l) { return l * 8u; }

// The stats words: the Adam step at [1], and the current batch's sample
// weights summed per block of the weight-sum kernel (at most kWeightBlocks
// blocks, their count at [2]), added in block order where used, so a
// training run gives the same network every time.
constexpr uint32_t kWeightSums = 16u, kWeightBlocks = 64u, kStatsWords = kWeightSums + kWeightBlocks;

// The network's device memory: master parameters, Adam moments, FP16
// fragments for the forward and backward products, bounds and step state.
struct NetBuffers {
    float *params;       // kParams (trained)
    float *m, *v;        // Adam moments
    _Float16 *fwd;       // kBlocks * 64 lanes * 8
    _Float16 *bwd;       // kBlocks * 64 lanes * 8
    float *ema;          // kParams: exponential moving average of the trained parameters (not bias-corrected)
    float *infer;        // kParams: the parameters inference uses (the corrected average)
    _Float16 *infer_fwd; // their forward fragments
    float *bounds;       // lo xyz, 1 / extent xyz
    float *stats;        // [1] Adam step, [2] blocks of the weight sum, [16 + b] block b's sum (kStatsWords)
};

// Lane group and element of position p (0..31) of a K = 32 block: two D
// tiles side by side.
__host__ __device__ inline void k_slot(uint32_t p, uint32_t &g, uint32_t &e) {
    g = (p % 16u) / 4u;
    e = p % 4u + (p >= 16u ? 4u : 0u);
}
// Index of W_l[in][out] in the forward fragments (A = W^T: row out, K in).
__host__ __device__ inline uint32_t fwd_index(uint32_t l, uint32_t in, uint32_t out) {
    uint32_t m = out / 16u, i = out % 16u, kb = in / 32u, g, e;
    k_slot(in % 32u, g, e);
    return ((block_base(l) + m * 2u + kb) * 64u + i + 16u * g) * 8u + e;
}
// Index of W_l[in][out] in the backward fragments (A = W: row in, K out).
__host__ __device__ inline uint32_t bwd_index(uint32_t l, uint32_t in, uint32_t out) {
    uint32_t blocks = layer_out(l) / 32u;
    uint32_t m = in / 16u, i = in % 16u, kb = out / 32u, g, e;
    k_slot(out % 32u, g, e);
    return ((block_base(l) + m * blocks + kb) * 64u + i + 16u * g) * 8u + e;
}

__device__ __forceinline__ half8_t load_block(const _Float16 *frags, uint32_t block, uint32_t lane) {
    return *reinterpret_cast<const half8_t *>(frags + ((size_t)block * 64u + lane) * 8u);
}
// Two D tiles' rows as one K = 32 B fragment.
__device__ __forceinline__ half8_t pair(half4_t lo, half4_t hi) {
    return half8_t{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}

__device__ __forceinline__ float4_t mfma32(half8_t a, half8_t b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// One-blob encoding (Mueller et al. 2019): a value in [0, 1] spread over
// `bins` Gaussian bumps of width 1 / bins.
__device__ __forceinline__ float blob(float u, uint32_t bin, uint32_t bins) {
    float d = (u - ((float)bin + 0.5f) / (float)bins) * (float)bins;
    return __expf(-0.5f * d * d);
}

// Octahedral map of a unit vector to [0, 1]^2.
__device__ __forceinline__ void octahedral_uv(float nx, float ny, float nz, float &u, float &v) {
    float s = fabsf(nx) + fabsf(ny) + fabsf(nz);
    s = s > 0.0f ? s : 1.0f;
    float x = nx / s, y = ny / s;
    if (nz < 0.0f) {
        float ox = (1.0f - fabsf(y)) * (x >= 0.0f ? 1.0f : -1.0f);
        float oy = (1.0f - fabsf(x)) * (y >= 0.0f ? 1.0f : -1.0f);
        x = ox, y = oy;
    }
    u = 0.5f * x + 0.5f, v = 0.5f * y + 0.5f;
}

// The input fragments of one point (B operand, K-step kk: features
// 16 kk + 4 g + r): kk 0..2 the position's x, y, z blobs, kk 3 the normal's
// octahedral u (features 0..7) and v (8..15) blobs.
__device__ __forceinline__ void encode(const float *bounds, float px, float py, float pz, float nx, float ny, float nz,
                                       uint32_t g, half4_t h[4]) {
    float u[3] = {(px - bounds[0]) * bounds[3], (py - bounds[1]) * bounds[4], (pz - bounds[2]) * bounds[5]};
    float ou, ov;
    octahedral_uv(nx, ny, nz, ou, ov);
#pragma unroll
    for (uint32_t kk = 0; kk < 3u; kk++) {
        float x = fminf(fmaxf(u[kk], 0.0f), 1.0f);
#pragma unroll
        for (uint32_t r = 0; r < 4u; r++) h[kk][r] = (_Float16)blob(x, 4u * g + r, kPosBins);
    }
#pragma unroll
    for (uint32_t r = 0; r < 4u; r++) {
        uint32_t f = 4u * g + r;
        h[3][r] = (_Float16)(f < 8u ? blob(ou, f, kNormalBins) : blob(ov, f - 8u, kNormalBins));
    }
}

// The whole encoding of one point into `row` (64 halves, features in the
// order encode's fragments take them: position x, y, z blobs, then the
// normal's u and v blobs).
__device__ __forceinline__ void encode_row(const float *bounds, float px, float py, float pz, float nx, float ny,
                                           float nz, _Float16 *row) {
    float u[3] = {(px - bounds[0]) * bounds[3], (py - bounds[1]) * bounds[4], (pz - bounds[2]) * bounds[5]};
    float ou, ov;
    octahedral_uv(nx, ny, nz, ou, ov);
#pragma unroll
    for (uint32_t kk = 0; kk < 3u; kk++) {
        float x = fminf(fmaxf(u[kk], 0.0f), 1.0f);
#pragma unroll
        for (uint32_t f = 0; f < kPosBins; f++) row[16u * kk + f] = (_Float16)blob(x, f, kPosBins);
    }
#pragma unroll
    for (uint32_t f = 0; f < 16u; f++) row[48u + f] = (_Float16)(f < 8u ? blob(ou, f, kNormalBins) : blob(ov, f - 8u, kNormalBins));
}

__device__ __forceinline__ half4_t to_half(float4_t x) {
    half4_t h;
#pragma unroll
    for (int r = 0; r < 4; r++) h[r] = (_Float16)x[r];
    return h;
}

// Z^T = W_l^T H^T + b for 16 points; `in` the four D tiles of H (the
// 64 inputs, two K = 32 blocks). Output tiles: layer_out(l) / 16.
__device__ __forceinline__ void layer_forward(const _Float16 *fwd, const float *params, uint32_t l,
                                              const half4_t in[4], uint32_t lane, float4_t out[4]) {
    uint32_t g = lane / 16u;
    uint32_t tiles = layer_out(l) / 16u;
    const half8_t b0 = pair(in[0], in[1]), b1 = pair(in[2], in[3]);
    for (uint32_t m = 0; m < tiles; m++) {
        float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mfma32(load_block(fwd, block_base(l) + m * 2u, lane), b0, acc);
        acc = mfma32(load_block(fwd, block_base(l) + m * 2u + 1u, lane), b1, acc);
        const float *b = params + bias_offset(l) + m * 16u + 4u * g;
#pragma unroll
        for (int r = 0; r < 4; r++) acc[r] += b[r];
        out[m] = acc;
    }
}

// ------------------------------------------------------------ the mixture

constexpr float kPi = 3.14159265358979323846f;
constexpr float kLogKappaLo = -4.0f, kLogKappaHi = 7.0f;

// Lobe k from the outputs o[5k .. 5k+4]: direction (normalized o[5k..5k+2]),
// concentration exp(clamp(o[5k+3])), weight softmax(o[5k+4]).
struct Mixture {
    float mu[kLobes][3];
    float kappa[kLobes];
    float w[kLobes];
};

__device__ __forceinline__ Mixture mixture(const float *o) {
    Mixture m;
    float top = -3.0e38f;
    for (uint32_t k = 0; k < kLobes; k++) top = fmaxf(top, o[5 * k + 4]);
    float total = 0.0f;
    for (uint32_t k = 0; k < kLobes; k++) {
        float x = o[5 * k], y = o[5 * k + 1], z = o[5 * k + 2];
        float len = sqrtf(x * x + y * y + z * z);
        if (len < 1e-6f) {
            x = 0.0f, y = 0.0f, z = 1.0f, len = 1.0f;
        }
        m.mu[k][0] = x / len, m.mu[k][1] = y / len, m.mu[k][2] = z / len;
        m.kappa[k] = __expf(fminf(fmaxf(o[5 * k + 3], kLogKappaLo), kLogKappaHi));
        m.w[k] = __expf(o[5 * k + 4] - top);
        total += m.w[k];
    }
    for (uint32_t k = 0; k < kLobes; k++) m.w[k] /= total;
    return m;
}

// vMF density on the sphere: kappa / (2 pi (1 - e^-2kappa)) e^(kappa (mu.d - 1)).
__device__ __forceinline__ float vmf(float kappa, float cos_mu) {
    return kappa / (2.0f * kPi * -expm1f(-2.0f * kappa)) * __expf(kappa * (cos_mu - 1.0f));
}

}  // namespace net
}  // namespace lr
