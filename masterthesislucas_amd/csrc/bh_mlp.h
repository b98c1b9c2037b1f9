// Learned far-field correction: a 10 -> 64 -> 64 -> 2 SiLU MLP that predicts, for a cell the
// opening-angle test accepts, the residual between the exact sum over the cell's particles and
// the cell's monopole, seen from the sink.
//
// Inputs (all O(1) by construction), with d = com - p, r = |d|, l = cell side, M = cell mass,
// S = mass-weighted second moments about the centre of mass, n = particle count:
//   f0, f1 = d / r          direction of the offset
//   f2     = l / r          offset normalised by the cell size (inverse), < theta when accepted
//   f3..f5 = Sxx, Sxy, Syy / (M l^2)
//   f6, f7 = cos 2phi, sin 2phi of the offset direction
//   f8     = log2(n) / 16
//   f9     = (l / r)^2
// Output y (2 floats): the residual in units of M l^2 / r^4, so the acceleration correction is
//   G * y * M l^2 / r^4.
//
// Parameters (FP32 master copy; W row-major [in x out]): W1[10x64] b1[64] W2[64x64] b2[64]
// W3[64x2] b3[2]. Inference uses BF16 copies laid out for the MFMA operands (see
// MlpBf16 below) with FP32 accumulation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bh {

constexpr int kMlpIn = 10;
constexpr int kMlpH = 64;
constexpr int kMlpOut = 2;
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpIn * kMlpH;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpH;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpH * kMlpH;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpH;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpH * kMlpOut;
constexpr int kMlpParams = kMlpOffB3 + kMlpOut;  // 4994
constexpr int kMlpInPad = 16;                    // MFMA K of the first layer; slot 10 holds 1

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// BF16 weights arranged per MFMA operand: for v_mfma_f32_16x16x16_bf16, lane l holds
// A[row = l % 16][k = 4 (l / 16) + 0..3] and B[k = 4 (l / 16) + 0..3][col = l % 16]; the result
// tile holds D[row = 4 (l / 16) + e][col = l % 16] in element e.
// Every layer computes the transposed product H^T = W^T X^T (rows = units, columns = samples),
// so a layer's result tile (unit block u, sample block s) is, element for element, the B
// operand of the next layer for K block u and column block s.
//   w1[u][l]     : A tile of W1^T for unit block u (b1 folded in as input row 10)
//   w2[u][k][l]  : A tile of W2^T for unit block u, K block k
//   w3[k][l]     : A tile of W3^T (rows 0, 1 real, rows 2..15 zero), K block k
//   b2, b3       : biases (FP32) used to initialise the accumulators
struct MlpBf16 {
    bf16x4 w1[4][64];
    bf16x4 w2[4][4][64];
    bf16x4 w3[4][64];
    float b2[kMlpH];
    float b3[kMlpOut];
};

__device__ __forceinline__ float mlp_silu(float v) {
    return v * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * v));
}

typedef __bf16 bf16v4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ short f32_to_bf16(float v) { return __builtin_bit_cast(short, (__bf16)v); }
__device__ __forceinline__ bf16x4 pack_bf16x4(float a, float b, float c, float d) {
    bf16v4 v;
    v[0] = (__bf16)a;
    v[1] = (__bf16)b;
    v[2] = (__bf16)c;
    v[3] = (__bf16)d;
    return __builtin_bit_cast(bf16x4, v);
}

// Features of one accepted (sink, cell) pair; returns the output scale M l^2 / r^4.
__device__ __forceinline__ float mlp_features(float dx, float dy, float r2, float l2, float M,
                                              uint32_t count, float sxx, float sxy, float syy,
                                              float* f) {
    const float inv_r = __builtin_amdgcn_rsqf(r2);
    const float ux = dx * inv_r, uy = dy * inv_r;
    const float l = __builtin_sqrtf(l2);
    const float q = l * inv_r;
    const float inv_ml2 = __builtin_amdgcn_rcpf(M * l2);
    f[0] = ux;
    f[1] = uy;
    f[2] = q;
    f[3] = sxx * inv_ml2;
    f[4] = sxy * inv_ml2;
    f[5] = syy * inv_ml2;
    f[6] = ux * ux - uy * uy;
    f[7] = 2.0f * ux * uy;
    f[8] = __builtin_amdgcn_logf((float)count) * (1.0f / 16.0f);  // log2
    f[9] = q * q;
    const float inv_r2 = inv_r * inv_r;
    return M * l2 * inv_r2 * inv_r2;
}

// One wave evaluates the MLP on 64 samples whose BF16 inputs sit in LDS as xs[64][16]
// (row = sample, slot 10 = 1.0, slots 11..15 = 0). Returns, in lanes 0..15 of each sample
// block s (s = 0..3), y0/y1 of sample 16 s + lane: out0[s], out1[s] (other lanes: garbage).
__device__ __forceinline__ void mlp_wave_infer(const MlpBf16* __restrict__ W,
                                               const short* __restrict__ xs, float out0[4],
                                               float out1[4]) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, q4 = lane >> 4;
    // layer 1: H1^T[64 x 64] = W1^T[64 x 16] X^T[16 x 64]
    bf16x4 h1[4][4];  // [unit block][sample block], B operands of layer 2
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const bf16x4 xb = *reinterpret_cast<const bf16x4*>(xs + (16 * s + r16) * kMlpInPad + 4 * q4);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W->w1[u][lane], xb, acc, 0, 0, 0);
            h1[u][s] = pack_bf16x4(mlp_silu(acc[0]), mlp_silu(acc[1]), mlp_silu(acc[2]),
                                   mlp_silu(acc[3]));
        }
    }
    // layer 2: H2^T = W2^T H1^T + b2
    bf16x4 h2[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        f32x4 bias;
#pragma unroll
        for (int e = 0; e < 4; ++e) bias[e] = W->b2[16 * u + 4 * q4 + e];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            f32x4 acc = bias;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W->w2[u][k][lane], h1[k][s], acc,
                                                                0, 0, 0);
            h2[u][s] = pack_bf16x4(mlp_silu(acc[0]), mlp_silu(acc[1]), mlp_silu(acc[2]),
                                   mlp_silu(acc[3]));
        }
    }
    // layer 3: Y^T[16 (2 real) x 64] = W3^T H2^T + b3
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        f32x4 acc;
        acc[0] = q4 == 0 ? W->b3[0] : 0.f;
        acc[1] = q4 == 0 ? W->b3[1] : 0.f;
        acc[2] = 0.f;
        acc[3] = 0.f;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W->w3[k][lane], h2[k][s], acc, 0, 0, 0);
        out0[s] = acc[0];  // row 0 (y0) of column 16 s + lane, valid in lanes 0..15
        out1[s] = acc[1];
    }
}

}  // namespace bh
