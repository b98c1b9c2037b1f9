// Learned far-field correction: a cell network and the correction it gives every pair.
//
// Per cell (mass M, centre of mass c, side l, n particles; s = x - c over its particles), once
// per build, a 10 -> 64 -> 64 -> 14 SiLU MLP maps the cell features
//   f0..f2  sum m s s^T / (M l^2)           (xx, xy, yy; mass-weighted second moments)
//   f3..f6  sum m s s s / (M l^3)           (xxx, xxy, xyy, yyy; 0 for cells of more than
//                                            kLearnT3Max particles)
//   f7, f8  (c - box centre) / l
//   f9      log2(n) / 16
// to 7 complex coefficients c_j. For an accepted (sink p, cell) pair with d = c - p, r = |d|,
// w = d / r (complex), q = l / r, the correction to the monopole acceleration (before G) is
//   (M l^2 / r^4) sum_j c_j b_j,   b = (w, conj w, w^3, q, q w^2, q conj(w)^2, q w^4):
// the angular forms of the quadrupole (q^0) and octupole (q^1) terms of the expansion of
// (d + s)/|d + s|^3 about s = 0, so the network's offset inputs (w, q) enter through this basis
// and its cell inputs through the coefficients. With c_1 = 0.75 tr S, c_2 = 0.375 Q,
// c_3 = 1.875 conj(Q) (Q = S_xx - S_yy + 2i S_xy, S normalised as f0..f2) the q^0 part is the
// exact quadrupole term; training on FP64 labels lets the network do better than that.
//
// Parameters (FP32 master copy; W row-major [in x out]): W1[10x64] b1[64] W2[64x64] b2[64]
// W3[64x14] b3[14]. Output j pairs (2j, 2j+1) are (Re, Im) of c_{j+1}.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bh {

constexpr int kMlpIn = 10;           // cell features
constexpr int kMlpH = 64;
constexpr int kLearnTerms = 7;       // complex basis functions
constexpr int kMlpOut = 2 * kLearnTerms;
constexpr int kLearnSample = kMlpIn + 3;  // data set row: cell features, w_x, w_y, q
constexpr int kCoefStride = 16;      // floats per node in the walk's coefficient records
constexpr uint32_t kLearnT3Max = 65536;  // third moments of cells up to this many particles
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpIn * kMlpH;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpH;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpH * kMlpH;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpH;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpH * kMlpOut;
constexpr int kMlpParams = kMlpOffB3 + kMlpOut;  // 5774
constexpr int kMlpInPad = 16;                    // K of the first layer; slot 10 holds 1

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// BF16 weights arranged per MFMA operand. Lane l (r16 = l % 16, q4 = l / 16):
//   v_mfma_f32_16x16x16_bf16 (layer 1, K = 16): A[row r16][k = 4 q4 + i], B[k = 4 q4 + i][col r16]
//   v_mfma_f32_16x16x32_bf16 (layers 2, 3, K = 32): A[row r16][k = 8 q4 + j], B[k = 8 q4 + j]
//   result tile D[row = 4 q4 + e][col r16] in element e (both forms).
// Each layer computes the transposed product (rows = units, columns = samples). Within a K = 32
// chunk c of the hidden units, logical k = 8 q4 + j is the physical unit 32 c + 4 q4 + j for
// j < 4 and 32 c + 16 + 4 q4 + (j - 4) for j >= 4: the four result elements a lane holds in unit
// blocks 2c and 2c + 1 are then, in order, its B operand for chunk c of the next layer.
//   w1[u][l]     : A tile of W1^T for unit block u (b1 folded in as input row 10)
//   w2[u][c][l]  : A tile of W2^T for output unit block u and K chunk c
//   w3[c][l]     : A tile of W3^T (rows 0..13 real, 14, 15 zero) for K chunk c
//   b2, b3       : biases (FP32) that initialise the accumulators
struct MlpBf16 {
    bf16x4 w1[4][64];
    bf16x8 w2[4][2][64];
    bf16x8 w3[2][64];
    float b2[kMlpH];
    float b3[16];
};

// physical hidden unit of logical K index k (0..31) in chunk c
__host__ __device__ constexpr int mlp_unit_of(int c, int k) {
    return 32 * c + ((k & 7) < 4 ? 4 * (k >> 3) + (k & 7) : 16 + 4 * (k >> 3) + (k & 7) - 4);
}

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
__device__ __forceinline__ bf16x8 silu_pack8(f32x4 lo, f32x4 hi) {
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        v[e] = (__bf16)mlp_silu(lo[e]);
        v[4 + e] = (__bf16)mlp_silu(hi[e]);
    }
    return v;
}

// One wave evaluates the network on 64 samples whose BF16 inputs sit in LDS as xs[64][16]
// (row = sample, slot 10 = 1.0, slots 11..15 = 0). out[s] holds, in lane l, outputs
// 4 (l / 16) + e (e = 0..3; 14 and 15 are zero) of sample 16 s + l % 16.
__device__ __forceinline__ void mlp_wave_infer(const MlpBf16* __restrict__ W,
                                               const short* __restrict__ xs, f32x4 out[4]) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, q4 = lane >> 4;
    // layer 1: H1^T[64 x 64] = W1^T[64 x 16] X^T[16 x 64], kept as f32 tiles [unit block][sample]
    f32x4 z1[4][4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const bf16x4 xb = *reinterpret_cast<const bf16x4*>(xs + (16 * s + r16) * kMlpInPad + 4 * q4);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
            z1[u][s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W->w1[u][lane], xb, zero, 0, 0, 0);
        }
    }
    bf16x8 h1[2][4];  // [K chunk][sample block]: B operands of layer 2
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int s = 0; s < 4; ++s) h1[c][s] = silu_pack8(z1[2 * c][s], z1[2 * c + 1][s]);
    // layer 2: H2^T = W2^T H1^T + b2
    f32x4 z2[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        f32x4 bias;
#pragma unroll
        for (int e = 0; e < 4; ++e) bias[e] = W->b2[16 * u + 4 * q4 + e];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            f32x4 acc = bias;
#pragma unroll
            for (int c = 0; c < 2; ++c)
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W->w2[u][c][lane], h1[c][s], acc, 0, 0, 0);
            z2[u][s] = acc;
        }
    }
    bf16x8 h2[2][4];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int s = 0; s < 4; ++s) h2[c][s] = silu_pack8(z2[2 * c][s], z2[2 * c + 1][s]);
    // layer 3: Y^T[16 (14 real) x 64] = W3^T H2^T + b3
    f32x4 bias3;
#pragma unroll
    for (int e = 0; e < 4; ++e) bias3[e] = W->b3[4 * q4 + e];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        f32x4 acc = bias3;
#pragma unroll
        for (int c = 0; c < 2; ++c)
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W->w3[c][lane], h2[c][s], acc, 0, 0, 0);
        out[s] = acc;
    }
}

// The walk's coefficient record (kCoefStride floats) from the 14 network outputs y (Re, Im of
// c_1..c_7) and the cell's M l^2 and side l: [A1, B1, C1, D1, a3, b3, a4, b4, A5, B5, C5, D5,
// a7, b7, M l^2, l] with A1 = a1 + a2, B1 = b2 - b1, C1 = a1 - a2, D1 = b1 + b2
// (c_1 w + c_2 conj w) and A5..D5 the same for c_5 w^2 + c_6 conj(w)^2.
__host__ __device__ inline void learned_record(const float* y, float ml2, float l, float* r) {
    r[0] = y[0] + y[2];
    r[1] = y[3] - y[1];
    r[2] = y[0] - y[2];
    r[3] = y[1] + y[3];
    r[4] = y[4];
    r[5] = y[5];
    r[6] = y[6];
    r[7] = y[7];
    r[8] = y[8] + y[10];
    r[9] = y[11] - y[9];
    r[10] = y[8] - y[10];
    r[11] = y[9] + y[11];
    r[12] = y[12];
    r[13] = y[13];
    r[14] = ml2;
    r[15] = l;
}

// Correction (before G) of one accepted pair: d = (dx, dy) = c - p, d2 = |d|^2, rec the cell's
// coefficient record (wave-uniform; M l^2 and l in its last two words). TERMS = 7: the whole
// basis; TERMS = 3: its quadrupole part c_1 w + c_2 conj w + c_3 w^3 only (a network trained
// with three terms).
template <int TERMS>
__device__ __forceinline__ void learned_correction(float dx, float dy, float d2, const float* rec,
                                                   float& cx, float& cy) {
    const float ml2 = rec[14], l = rec[15];
    static_assert(TERMS == 3 || TERMS == 7, "3 or 7 basis terms");
    const float inv_r = __builtin_amdgcn_rsqf(d2);
    const float ux = dx * inv_r, uy = dy * inv_r;
    const float w2r = fmaf(ux, ux, -uy * uy), w2i = 2.0f * ux * uy;
    const float w3r = fmaf(w2r, ux, -w2i * uy), w3i = fmaf(w2r, uy, w2i * ux);
    // q^0 part: c_1 w + c_2 conj w + c_3 w^3
    float re = fmaf(rec[0], ux, fmaf(rec[1], uy, fmaf(rec[4], w3r, -rec[5] * w3i)));
    float im = fmaf(rec[2], uy, fmaf(rec[3], ux, fmaf(rec[4], w3i, rec[5] * w3r)));
    if constexpr (TERMS == 7) {
        // q^1 part: c_4 + c_5 w^2 + c_6 conj(w)^2 + c_7 w^4, times q = l / r
        const float q = l * inv_r;
        const float w4r = fmaf(w2r, w2r, -w2i * w2i), w4i = 2.0f * w2r * w2i;
        float tr = fmaf(rec[8], w2r, fmaf(rec[9], w2i, rec[6]));
        tr = fmaf(rec[12], w4r, fmaf(-rec[13], w4i, tr));
        float ti = fmaf(rec[10], w2i, fmaf(rec[11], w2r, rec[7]));
        ti = fmaf(rec[12], w4i, fmaf(rec[13], w4r, ti));
        re = fmaf(q, tr, re);
        im = fmaf(q, ti, im);
    }
    const float inv_r2 = inv_r * inv_r;
    const float s = ml2 * inv_r2 * inv_r2;
    cx = s * re;
    cy = s * im;
}

}  // namespace bh
