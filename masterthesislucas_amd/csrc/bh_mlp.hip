// Learned far-field correction on gfx950: training data from the tree itself, hand-written
// forward / backward / Adam kernels (FP32), and BF16 MFMA inference (FP32 accumulation).
//
// Data set: for sampled sinks, the cells the opening-angle test accepts (within a band of
// l / r) are collected by a per-sink stackless walk over the pre-order tree, then each
// (sink, cell) pair is labelled by an FP64 sum over the cell's particles minus the FP64
// monopole, in units of M l^2 / r^4 (bh_mlp.h). Training minimises the mean squared error over
// all outputs with Adam; gradients are reduced per chunk of samples and then summed in a fixed
// order, so a training run is deterministic.
#include "bh_common.h"
#include "bh_launch.h"
#include "bh_mlp.h"

namespace bh {

static inline uint32_t div_up_u(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ---------------------------------------------------------------------------------------------
// parameters -> MFMA operand layout
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_mlp_prep_bf16(const float* __restrict__ p,
                                                      MlpBf16* __restrict__ W) {
    const int lane = threadIdx.x;
    const int r16 = lane & 15, q4 = lane >> 4;
    const float* W1 = p + kMlpOffW1;
    const float* b1 = p + kMlpOffB1;
    const float* W2 = p + kMlpOffW2;
    const float* W3 = p + kMlpOffW3;
    for (int u = 0; u < 4; ++u) {
        bf16x4 v;
        for (int i = 0; i < 4; ++i) {
            const int in = 4 * q4 + i, unit = 16 * u + r16;
            float w = 0.f;
            if (in < kMlpIn) w = W1[in * kMlpH + unit];
            else if (in == kMlpIn) w = b1[unit];  // constant-1 input carries the bias
            v[i] = f32_to_bf16(w);
        }
        W->w1[u][lane] = v;
    }
    for (int u = 0; u < 4; ++u)
        for (int c = 0; c < 2; ++c) {
            bf16x8 v;
            for (int j = 0; j < 8; ++j)
                v[j] = (__bf16)W2[mlp_unit_of(c, 8 * q4 + j) * kMlpH + 16 * u + r16];
            W->w2[u][c][lane] = v;
        }
    for (int c = 0; c < 2; ++c) {
        bf16x8 v;
        for (int j = 0; j < 8; ++j)
            v[j] = (__bf16)(r16 < kMlpOut ? W3[mlp_unit_of(c, 8 * q4 + j) * kMlpOut + r16] : 0.f);
        W->w3[c][lane] = v;
    }
    W->b2[lane] = p[kMlpOffB2 + lane];
    if (lane < 16) W->b3[lane] = lane < kMlpOut ? p[kMlpOffB3 + lane] : 0.f;
}

// stage the BF16 weights in LDS once per block
__device__ __forceinline__ void stage_weights(const MlpBf16* __restrict__ W, MlpBf16* ws) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(W);
    uint32_t* dst = reinterpret_cast<uint32_t*>(ws);
    for (uint32_t i = threadIdx.x; i < sizeof(MlpBf16) / 4; i += kBlock) dst[i] = src[i];
}

// one wave: features of 64 rows (stride `stride` floats, the first kMlpIn used) -> the 14
// outputs of each row in LDS (ys[64][16])
__device__ __forceinline__ void infer_rows(const MlpBf16* ws, const float* __restrict__ X,
                                           uint32_t base, uint32_t n, uint32_t stride, short* xs,
                                           float* ys) {
    const int lane = threadIdx.x & 63;
    const uint32_t i = base + lane;
    short* row = xs + lane * kMlpInPad;
#pragma unroll
    for (int c = 0; c < kMlpInPad; ++c) {
        float v = 0.f;
        if (c < kMlpIn) v = i < n ? X[(size_t)i * stride + c] : 0.f;
        else if (c == kMlpIn) v = 1.f;
        row[c] = f32_to_bf16(v);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    f32x4 out[4];
    mlp_wave_infer(ws, xs, out);
    const int r16 = lane & 15, q4 = lane >> 4;
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int e = 0; e < 4; ++e) ys[(16 * s + r16) * 16 + 4 * q4 + e] = out[s][e];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------------------------------------
// inference: BF16 MFMA (one wave per 64 samples) and the FP32 reference forward
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_mlp_infer_bf16(const float* __restrict__ X, uint32_t n,
                                                           const MlpBf16* __restrict__ W,
                                                           float* __restrict__ Y) {
    __shared__ short xs[kWavesPerBlock][64 * kMlpInPad];
    __shared__ float ys[kWavesPerBlock][64 * 16];
    __shared__ MlpBf16 ws;
    stage_weights(W, &ws);
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t base = (blockIdx.x * kWavesPerBlock + w) * 64u;
    infer_rows(&ws, X, base, n, kMlpIn, xs[w], ys[w]);
    const uint32_t i = base + lane;
    if (i < n)
        for (int o = 0; o < kMlpOut; ++o) Y[(size_t)i * kMlpOut + o] = ys[w][lane * 16 + o];
}

// per node: the network on its features -> the walk's coefficient record
__global__ __launch_bounds__(kBlock) void k_cell_coefs(const float* __restrict__ feat,
                                                       const Node* __restrict__ nodes,
                                                       const BuildParams* __restrict__ prm,
                                                       uint32_t cap_nodes,
                                                       const MlpBf16* __restrict__ W,
                                                       float4* __restrict__ coef) {
    __shared__ short xs[kWavesPerBlock][64 * kMlpInPad];
    __shared__ float ys[kWavesPerBlock][64 * 16];
    __shared__ MlpBf16 ws;
    const uint32_t nn = min(prm->n_nodes, cap_nodes);
    if (blockIdx.x * (64u * kWavesPerBlock) >= nn) return;  // whole block: past the last node
    stage_weights(W, &ws);
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t base = (blockIdx.x * kWavesPerBlock + w) * 64u;
    if (base >= nn) return;  // wave-uniform
    infer_rows(&ws, feat, base, nn, kMlpIn, xs[w], ys[w]);
    const uint32_t k = base + lane;
    if (k >= nn) return;
    // the cell side as the walk takes it (the FP32 root side scaled by 2^-level, exact) and
    // M l^2 with the walk's rounding
    const Node nd = nodes[k];
    const float l = ldexpf(prm->side, -(int)(nd.meta & kMetaLevelMask));
    float r[kCoefStride];
    learned_record(&ys[w][lane * 16], nd.mass * l * l, l, r);
    float4* dst = coef + (size_t)k * (kCoefStride / 4);
#pragma unroll
    for (int q = 0; q < kCoefStride / 4; ++q)
        dst[q] = make_float4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
}

__device__ __forceinline__ float silu_f(float v) { return v / (1.0f + __expf(-v)); }
__device__ __forceinline__ float dsilu_f(float v) {
    const float s = 1.0f / (1.0f + __expf(-v));
    return s * (1.0f + v * (1.0f - s));
}

__global__ __launch_bounds__(kBlock) void k_mlp_forward_f32(const float* __restrict__ X, uint32_t n,
                                                            const float* __restrict__ p,
                                                            float* __restrict__ Y) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float x[kMlpIn], a1[kMlpH], y[kMlpOut];
#pragma unroll
    for (int c = 0; c < kMlpIn; ++c) x[c] = X[(size_t)i * kMlpIn + c];
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) {
        float z = p[kMlpOffB1 + j];
#pragma unroll
        for (int c = 0; c < kMlpIn; ++c) z = fmaf(x[c], p[kMlpOffW1 + c * kMlpH + j], z);
        a1[j] = silu_f(z);
    }
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o) y[o] = p[kMlpOffB3 + o];
    for (int j = 0; j < kMlpH; ++j) {
        float z = p[kMlpOffB2 + j];
#pragma unroll
        for (int k = 0; k < kMlpH; ++k) z = fmaf(a1[k], p[kMlpOffW2 + k * kMlpH + j], z);
        const float a2 = silu_f(z);
#pragma unroll
        for (int o = 0; o < kMlpOut; ++o) y[o] = fmaf(a2, p[kMlpOffW3 + j * kMlpOut + o], y[o]);
    }
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o) Y[(size_t)i * kMlpOut + o] = y[o];
}

// basis of the pair offset (complex, interleaved re/im): w, conj w, w^3, q, q w^2, q conj(w)^2,
// q w^4
__device__ __forceinline__ void learned_basis_f(float ux, float uy, float q, float* b) {
    const float w2r = fmaf(ux, ux, -uy * uy), w2i = 2.0f * ux * uy;
    const float w3r = fmaf(w2r, ux, -w2i * uy), w3i = fmaf(w2r, uy, w2i * ux);
    const float w4r = fmaf(w2r, w2r, -w2i * w2i), w4i = 2.0f * w2r * w2i;
    b[0] = ux; b[1] = uy;
    b[2] = ux; b[3] = -uy;
    b[4] = w3r; b[5] = w3i;
    b[6] = q; b[7] = 0.f;
    b[8] = q * w2r; b[9] = q * w2i;
    b[10] = q * w2r; b[11] = -q * w2i;
    b[12] = q * w4r; b[13] = q * w4i;
}

// ---------------------------------------------------------------------------------------------
// training: per-sample forward + backward to the pre-activation gradients, written
// feature-major to scratch S (rows of length B): A1[64] DZ1[64] A2[64] DZ2[64] DY[14] X[10].
// A sample is a row of kLearnSample floats (10 cell features, w_x, w_y, q) and a label
// T[2]; its prediction is p = sum_j c_j b_j (complex) with c = the network's 14 outputs and b
// the basis of (w, q); loss = sum |p - T|^2 / (2 B).
// ---------------------------------------------------------------------------------------------
constexpr int kRowA1 = 0, kRowDZ1 = 64, kRowA2 = 128, kRowDZ2 = 192, kRowDY = 256;
constexpr int kRowX = kRowDY + kMlpOut;
constexpr int kScratchRows = kRowX + kMlpIn;

__global__ __launch_bounds__(kBlock) void k_mlp_fwdbwd(const float* __restrict__ X,
                                                       const float* __restrict__ T,
                                                       const uint32_t* __restrict__ order,
                                                       uint32_t B, const float* __restrict__ p,
                                                       float* __restrict__ S,
                                                       double* __restrict__ loss_part, int terms) {
    __shared__ double red[kWavesPerBlock];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    double sq = 0.0;
    if (i < B) {
        const uint32_t src = order ? order[i] : i;
        const float* row = X + (size_t)src * kLearnSample;
        float x[kMlpIn];
#pragma unroll
        for (int c = 0; c < kMlpIn; ++c) {
            x[c] = row[c];
            S[(size_t)(kRowX + c) * B + i] = x[c];
        }
        float a1[kMlpH];
#pragma unroll
        for (int j = 0; j < kMlpH; ++j) {
            float z = p[kMlpOffB1 + j];
#pragma unroll
            for (int c = 0; c < kMlpIn; ++c) z = fmaf(x[c], p[kMlpOffW1 + c * kMlpH + j], z);
            a1[j] = silu_f(z);
            S[(size_t)(kRowA1 + j) * B + i] = a1[j];
            S[(size_t)(kRowDZ1 + j) * B + i] = dsilu_f(z);  // silu'(z1), scaled below
        }
        float y[kMlpOut];
#pragma unroll
        for (int o = 0; o < kMlpOut; ++o) y[o] = p[kMlpOffB3 + o];
        for (int j = 0; j < kMlpH; ++j) {
            float z = p[kMlpOffB2 + j];
#pragma unroll
            for (int k = 0; k < kMlpH; ++k) z = fmaf(a1[k], p[kMlpOffW2 + k * kMlpH + j], z);
            const float a2 = silu_f(z);
            S[(size_t)(kRowA2 + j) * B + i] = a2;
            S[(size_t)(kRowDZ2 + j) * B + i] = dsilu_f(z);  // silu'(z2), scaled below
#pragma unroll
            for (int o = 0; o < kMlpOut; ++o) y[o] = fmaf(a2, p[kMlpOffW3 + j * kMlpOut + o], y[o]);
        }
        float bas[kMlpOut];
        learned_basis_f(row[kMlpIn], row[kMlpIn + 1], row[kMlpIn + 2], bas);
        // a network trained with fewer terms: the basis beyond them is zero (no prediction, no
        // gradient)
#pragma unroll
        for (int o = 0; o < kMlpOut; ++o)
            if (o >= 2 * terms) bas[o] = 0.f;
        float pr = 0.f, pi = 0.f;
#pragma unroll
        for (int j = 0; j < kLearnTerms; ++j) {
            pr = fmaf(y[2 * j], bas[2 * j], fmaf(-y[2 * j + 1], bas[2 * j + 1], pr));
            pi = fmaf(y[2 * j], bas[2 * j + 1], fmaf(y[2 * j + 1], bas[2 * j], pi));
        }
        const float ex = pr - T[(size_t)src * 2], ey = pi - T[(size_t)src * 2 + 1];
        sq = (double)ex * ex + (double)ey * ey;
        // loss = sum |e|^2 / (2 B): dL/dRe c_j = Re(conj(e) b_j) / B, dL/dIm c_j = Im(conj(b_j) e)/B
        const float inv_b = 1.0f / (float)B;
        float dy[kMlpOut];
#pragma unroll
        for (int j = 0; j < kLearnTerms; ++j) {
            const float br = bas[2 * j], bi = bas[2 * j + 1];
            dy[2 * j] = fmaf(ex, br, ey * bi) * inv_b;
            dy[2 * j + 1] = fmaf(ey, br, -ex * bi) * inv_b;
            S[(size_t)(kRowDY + 2 * j) * B + i] = dy[2 * j];
            S[(size_t)(kRowDY + 2 * j + 1) * B + i] = dy[2 * j + 1];
        }
        float dz2[kMlpH];
#pragma unroll
        for (int j = 0; j < kMlpH; ++j) {
            float da2 = 0.f;
#pragma unroll
            for (int o = 0; o < kMlpOut; ++o) da2 = fmaf(dy[o], p[kMlpOffW3 + j * kMlpOut + o], da2);
            dz2[j] = da2 * S[(size_t)(kRowDZ2 + j) * B + i];
            S[(size_t)(kRowDZ2 + j) * B + i] = dz2[j];
        }
        for (int k = 0; k < kMlpH; ++k) {
            float da1 = 0.f;
#pragma unroll
            for (int j = 0; j < kMlpH; ++j) da1 = fmaf(dz2[j], p[kMlpOffW2 + k * kMlpH + j], da1);
            S[(size_t)(kRowDZ1 + k) * B + i] *= da1;
        }
    }
    if (loss_part) {
        for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
        __syncthreads();
        if (threadIdx.x == 0) {
            double b = 0.0;
            for (int k = 0; k < kWavesPerBlock; ++k) b += red[k];
            loss_part[blockIdx.x] = b;
        }
    }
}

// gradient partials: block (param group g, sample chunk c); each thread owns parameters and
// sums row products over the chunk
constexpr int kGradChunk = 1024;

__device__ __forceinline__ void param_rows(int q, int& ra, int& rb) {
    // returns the scratch rows whose product summed over samples is dL/dparam q (rb = -1: bias)
    if (q < kMlpOffB1) { ra = kRowX + q / kMlpH; rb = kRowDZ1 + q % kMlpH; return; }
    if (q < kMlpOffW2) { ra = -1; rb = kRowDZ1 + (q - kMlpOffB1); return; }
    if (q < kMlpOffB2) { const int r = q - kMlpOffW2; ra = kRowA1 + r / kMlpH; rb = kRowDZ2 + r % kMlpH; return; }
    if (q < kMlpOffW3) { ra = -1; rb = kRowDZ2 + (q - kMlpOffB2); return; }
    if (q < kMlpOffB3) { const int r = q - kMlpOffW3; ra = kRowA2 + r / kMlpOut; rb = kRowDY + r % kMlpOut; return; }
    ra = -1;
    rb = kRowDY + (q - kMlpOffB3);
}

__global__ __launch_bounds__(kBlock) void k_mlp_wgrad(const float* __restrict__ S, uint32_t B,
                                                      float* __restrict__ partials) {
    const int q = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t c = blockIdx.y;
    if (q >= kMlpParams) return;
    int ra, rb;
    param_rows(q, ra, rb);
    const uint32_t s0 = c * kGradChunk, s1 = min(B, s0 + kGradChunk);
    const float* pb = S + (size_t)rb * B;
    float acc = 0.f;
    if (ra < 0) {
        for (uint32_t s = s0; s < s1; ++s) acc += pb[s];
    } else {
        const float* pa = S + (size_t)ra * B;
        for (uint32_t s = s0; s < s1; ++s) acc = fmaf(pa[s], pb[s], acc);
    }
    partials[(size_t)c * kMlpParams + q] = acc;
}

// sum the chunk partials in order; optionally export the gradient; Adam update
__global__ __launch_bounds__(kBlock) void k_mlp_adam(float* __restrict__ p,
                                                     const float* __restrict__ partials,
                                                     uint32_t nchunks, float* __restrict__ m1,
                                                     float* __restrict__ m2, float lr, float beta1,
                                                     float beta2, float eps, int step,
                                                     float* __restrict__ grad_out, int apply) {
    const int q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= kMlpParams) return;
    float g = 0.f;
    for (uint32_t c = 0; c < nchunks; ++c) g += partials[(size_t)c * kMlpParams + q];
    if (grad_out) grad_out[q] = g;
    if (!apply) return;
    const float a = m1[q] = beta1 * m1[q] + (1.f - beta1) * g;
    const float b = m2[q] = beta2 * m2[q] + (1.f - beta2) * g * g;
    const float bc1 = 1.f - __powf(beta1, (float)step), bc2 = 1.f - __powf(beta2, (float)step);
    p[q] -= lr * (a / bc1) / (sqrtf(b / bc2) + eps);
}

// Adam with a given gradient (host-provided), same update as above
__global__ __launch_bounds__(kBlock) void k_mlp_adam_grad(float* __restrict__ p,
                                                          const float* __restrict__ g_in,
                                                          float* __restrict__ m1,
                                                          float* __restrict__ m2, float lr,
                                                          float beta1, float beta2, float eps,
                                                          int step) {
    const int q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= kMlpParams) return;
    const float g = g_in[q];
    const float a = m1[q] = beta1 * m1[q] + (1.f - beta1) * g;
    const float b = m2[q] = beta2 * m2[q] + (1.f - beta2) * g * g;
    const float bc1 = 1.f - __powf(beta1, (float)step), bc2 = 1.f - __powf(beta2, (float)step);
    p[q] -= lr * (a / bc1) / (sqrtf(b / bc2) + eps);
}

// ---------------------------------------------------------------------------------------------
// data set from the tree
// ---------------------------------------------------------------------------------------------
// Per sample sink (sorted index), a stackless pre-order walk with the traversal's own FP32
// acceptance test; accepted cells with d^2 <= l^2 / band^2 (l / r >= band) are recorded.
__global__ __launch_bounds__(kBlock) void k_collect_pairs(const float4* __restrict__ pos4,
                                                          const Node* __restrict__ nodes,
                                                          const BuildParams* __restrict__ prm,
                                                          const uint32_t* __restrict__ sinks,
                                                          uint32_t ns, float inv_theta2,
                                                          float inv_band2, uint32_t cap,
                                                          uint32_t* __restrict__ pairs,
                                                          uint32_t* __restrict__ counts) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= ns) return;
    const uint32_t nn = prm->n_nodes;
    const float4 me = pos4[sinks[t]];
    uint32_t k = 0, c = 0;
    while (k < nn) {
        const Node nd = nodes[k];
        const float dx = nd.cx - me.x, dy = nd.cy - me.y;
        const float d2 = fmaf(dx, dx, dy * dy);
        const uint32_t next = max(nd.skip, k + 1);
        if (d2 > nd.l2 * inv_theta2) {
            if ((inv_band2 == 0.f || !(d2 > nd.l2 * inv_band2)) && c < cap) pairs[(size_t)t * cap + c++] = k;
            k = next;
        } else if (nd.meta & kMetaPP) {
            k = next;
        } else {
            k = k + 1;
        }
    }
    counts[t] = c;
}

// One wave per pair (slot t, c-th cell): FP64 sum over the cell's particles (sink excluded by
// index) minus the FP64 monopole, in units of M l^2 / r^4; the sample row holds the cell's
// features and the pair's offset features w = d / r, q = l / r (d = c - p).
__global__ __launch_bounds__(kBlock) void k_label_pairs(const float4* __restrict__ pos4,
                                                        const Node* __restrict__ nodes,
                                                        const float* __restrict__ feat,
                                                        const BuildParams* __restrict__ prm,
                                                        const uint32_t* __restrict__ sinks,
                                                        const uint32_t* __restrict__ pairs,
                                                        const uint32_t* __restrict__ counts,
                                                        const uint32_t* __restrict__ offsets,
                                                        uint32_t ns, uint32_t cap, double eps2,
                                                        float* __restrict__ Xo,
                                                        float* __restrict__ To, uint32_t total) {
    const int lane = threadIdx.x & 63;
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (gw >= total) return;
    // find the slot whose range holds gw (offsets is an exclusive scan of counts)
    uint32_t lo = 0, hi = ns;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (offsets[mid] <= gw) lo = mid; else hi = mid;
    }
    const uint32_t t = lo, c = gw - offsets[t];
    const uint32_t sink = sinks[t];
    const uint32_t k = pairs[(size_t)t * cap + c];
    const Node nd = nodes[k];
    const float4 me = pos4[sink];
    const double px = me.x, py = me.y;
    double sx = 0.0, sy = 0.0;
    for (uint32_t j = nd.first + lane; j < nd.first + nd.count; j += 64) {
        if (j == sink) continue;
        const float4 q = pos4[j];
        const double dx = (double)q.x - px, dy = (double)q.y - py;
        const double r2 = dx * dx + dy * dy + eps2;
        const double inv = 1.0 / sqrt(r2);
        const double s = (double)q.z * inv * inv * inv;
        sx += s * dx;
        sy += s * dy;
    }
    for (int o = 32; o > 0; o >>= 1) {
        sx += __shfl_xor(sx, o, 64);
        sy += __shfl_xor(sy, o, 64);
    }
    if (lane != 0) return;
    const double dx = (double)nd.cx - px, dy = (double)nd.cy - py;
    const double r2m = dx * dx + dy * dy;
    const double invm = 1.0 / sqrt(r2m + eps2);
    const double sm = (double)nd.mass * invm * invm * invm;
    // the cell side as the walk takes it: the FP32 root side scaled by 2^-level (exact)
    const double l = ldexp((double)prm->side, -(int)(nd.meta & kMetaLevelMask));
    const double rm = sqrt(r2m);
    const double scale = (double)nd.mass * l * l / (r2m * r2m);
    float* row = Xo + (size_t)gw * kLearnSample;
#pragma unroll
    for (int q = 0; q < kMlpIn; ++q) row[q] = feat[(size_t)k * kMlpIn + q];
    row[kMlpIn] = (float)(dx / rm);
    row[kMlpIn + 1] = (float)(dy / rm);
    row[kMlpIn + 2] = (float)(l / rm);
    To[(size_t)gw * 2] = (float)((sx - sm * dx) / scale);
    To[(size_t)gw * 2 + 1] = (float)((sy - sm * dy) / scale);
}

// even (x) and odd (y) bits of a Morton prefix, compacted
__device__ __forceinline__ uint64_t compact_even(uint64_t v) {
    v &= 0x5555555555555555ull;
    v = (v | (v >> 1)) & 0x3333333333333333ull;
    v = (v | (v >> 2)) & 0x0F0F0F0F0F0F0F0Full;
    v = (v | (v >> 4)) & 0x00FF00FF00FF00FFull;
    v = (v | (v >> 8)) & 0x0000FFFF0000FFFFull;
    v = (v | (v >> 16)) & 0x00000000FFFFFFFFull;
    return v;
}

__device__ __forceinline__ double f64_of_bits(uint64_t b) { return __longlong_as_double((long long)b); }

// third moments sum m s s s about (cx, cy) over particles [f, e), `stride` threads apart from
// `t0`, kFeatLoads loads in flight per thread (pos4 is padded past n; the tail is masked)
constexpr uint32_t kFeatLoads = 4;
__device__ __forceinline__ void third_moments(const float4* __restrict__ pos4, uint32_t f,
                                              uint32_t e, uint32_t t0, uint32_t stride, float cx,
                                              float cy, float* t) {
    for (uint32_t j0 = f + t0; j0 < e; j0 += kFeatLoads * stride) {
        float4 q[kFeatLoads];
#pragma unroll
        for (uint32_t u = 0; u < kFeatLoads; ++u) q[u] = pos4[min(j0 + u * stride, e - 1)];
#pragma unroll
        for (uint32_t u = 0; u < kFeatLoads; ++u) {
            const float m = j0 + u * stride < e ? q[u].z : 0.f;
            const float x = q[u].x - cx, y = q[u].y - cy;
            const float mx = m * x, my = m * y;
            t[0] = fmaf(mx * x, x, t[0]);
            t[1] = fmaf(mx * x, y, t[1]);
            t[2] = fmaf(mx * y, y, t[2]);
            t[3] = fmaf(my * y, y, t[3]);
        }
    }
}

// wave-aggregated append of the lanes with `want` set to list[1 + ...] (list[0] counts)
__device__ __forceinline__ void list_append(uint32_t* __restrict__ list, uint32_t cap, bool want,
                                            uint32_t v) {
    const uint64_t m = __ballot(want);
    if (!m) return;
    const int lane = threadIdx.x & 63;
    const int leader = __builtin_ctzll(m);
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&list[0], (uint32_t)__popcll(m));
    base = __shfl(base, leader, kWave);
    const uint32_t slot = base + (uint32_t)__popcll(m & lane_mask_lt());
    if (want && slot < cap) list[1 + slot] = v;
}

__device__ __forceinline__ float cell_side_f(const BuildParams* __restrict__ prm, uint32_t meta) {
    return (float)ldexp(f64_of_bits(prm->side64), -(int)(meta & kMetaLevelMask));
}

// One thread per node: the cell features (see bh_mlp.h). Second moments come from the upward
// pass (m2); third moments are summed about the node's centre of mass over its particles, here
// for cells of at most kFeatThread particles; larger cells are listed for one wave each
// (k_cell_t3_wave, up to kFeatWave particles) or one block each (k_cell_t3_big, up to
// kLearnT3Max). The box centre follows from the Morton prefix of the node's first particle (B
// code bits per axis) and the root square, in FP64.
constexpr uint32_t kFeatThread = 64;
constexpr uint32_t kFeatWave = 2048;
template <typename K>
__global__ __launch_bounds__(kBlock) void k_cell_features(const float4* __restrict__ pos4,
                                                          const Node* __restrict__ nodes,
                                                          const float4* __restrict__ m2,
                                                          const K* __restrict__ codes, int B,
                                                          const BuildParams* __restrict__ prm,
                                                          uint32_t cap_nodes,
                                                          float* __restrict__ feat,
                                                          uint32_t* __restrict__ wlist,
                                                          uint32_t wcap,
                                                          uint32_t* __restrict__ blist,
                                                          uint32_t bcap) {
    const uint32_t nn = min(prm->n_nodes, cap_nodes);
    if (blockIdx.x * kBlock >= nn) return;  // whole block
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    const bool live = k < nn;
    const Node nd = nodes[live ? k : nn - 1];
    list_append(wlist, wcap, live && nd.count > kFeatThread && nd.count <= kFeatWave, k);
    list_append(blist, bcap, live && nd.count > kFeatWave && nd.count <= kLearnT3Max, k);
    if (!live) return;
    float t[4] = {0.f, 0.f, 0.f, 0.f};
    if (nd.count <= kFeatThread) third_moments(pos4, nd.first, nd.first + nd.count, 0, 1, nd.cx, nd.cy, t);
    const int L = (int)(nd.meta & kMetaLevelMask);
    const uint64_t pre = L > 0 ? (uint64_t)codes[nd.first] >> (2 * (B - L)) : 0ull;
    const double l64 = ldexp(f64_of_bits(prm->side64), -L);
    const double bx = f64_of_bits(prm->minx64) + ((double)compact_even(pre) + 0.5) * l64;
    const double by = f64_of_bits(prm->miny64) + ((double)compact_even(pre >> 1) + 0.5) * l64;
    const float l = (float)l64;
    const float4 s2 = m2[k];
    const float a = 1.0f / (nd.mass * l * l), b = a / l;
    float* o = feat + (size_t)k * kMlpIn;
    o[0] = s2.x * a;
    o[1] = s2.y * a;
    o[2] = s2.z * a;
    o[3] = t[0] * b;
    o[4] = t[1] * b;
    o[5] = t[2] * b;
    o[6] = t[3] * b;
    o[7] = (float)(((double)nd.cx - bx) / l64);
    o[8] = (float)(((double)nd.cy - by) / l64);
    o[9] = (float)(log2((double)nd.count) / 16.0);
}

// Third moments of the cells listed for one wave each (grid-stride over the list).
__global__ __launch_bounds__(kBlock) void k_cell_t3_wave(const float4* __restrict__ pos4,
                                                         const Node* __restrict__ nodes,
                                                         const BuildParams* __restrict__ prm,
                                                         const uint32_t* __restrict__ wlist,
                                                         uint32_t wcap,
                                                         float* __restrict__ feat) {
    const uint32_t nw = min(wlist[0], wcap);
    const int lane = threadIdx.x & 63;
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    for (uint32_t e = gw; e < nw; e += gridDim.x * kWavesPerBlock) {
        const uint32_t k = wlist[1 + e];
        const Node nd = nodes[k];
        float t[4] = {0.f, 0.f, 0.f, 0.f};
        third_moments(pos4, nd.first, nd.first + nd.count, lane, kWave, nd.cx, nd.cy, t);
#pragma unroll
        for (int i = 0; i < 4; ++i) t[i] = wave_sum_f(t[i]);
        if (lane < 4) {
            const float l = cell_side_f(prm, nd.meta);
            const float b = 1.0f / (nd.mass * l * l) / l;
            const float v = lane == 0 ? t[0] : lane == 1 ? t[1] : lane == 2 ? t[2] : t[3];
            feat[(size_t)k * kMlpIn + 3 + lane] = v * b;
        }
    }
}

// Third moments of the listed large cells: one block of kBigThreads per cell (grid-stride over
// the list), reduced in a fixed order.
constexpr int kBigThreads = 1024;
__global__ __launch_bounds__(kBigThreads) void k_cell_t3_big(const float4* __restrict__ pos4,
                                                             const Node* __restrict__ nodes,
                                                             const BuildParams* __restrict__ prm,
                                                             const uint32_t* __restrict__ big,
                                                             uint32_t big_cap,
                                                             float* __restrict__ feat) {
    __shared__ float red[kBigThreads / kWave][4];
    const uint32_t nb = min(big[0], big_cap);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint32_t e = blockIdx.x; e < nb; e += gridDim.x) {
        const uint32_t k = big[1 + e];
        const Node nd = nodes[k];
        float t[4] = {0.f, 0.f, 0.f, 0.f};
        third_moments(pos4, nd.first, nd.first + nd.count, threadIdx.x, kBigThreads, nd.cx,
                      nd.cy, t);
#pragma unroll
        for (int i = 0; i < 4; ++i) t[i] = wave_sum_f(t[i]);
        if (lane == 0)
#pragma unroll
            for (int i = 0; i < 4; ++i) red[w][i] = t[i];
        __syncthreads();
        if (threadIdx.x < 4) {
            float v = 0.f;
            for (int q = 0; q < kBigThreads / kWave; ++q) v += red[q][threadIdx.x];
            const float l = cell_side_f(prm, nd.meta);
            const float b = 1.0f / (nd.mass * l * l) / l;
            feat[(size_t)k * kMlpIn + 3 + threadIdx.x] = v * b;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------------
void launch_mlp_prep(hipStream_t s, const float* params, MlpBf16* w) {
    k_mlp_prep_bf16<<<1, 64, 0, s>>>(params, w);
}
void launch_mlp_infer_bf16(hipStream_t s, const float* X, uint32_t n, const MlpBf16* w, float* Y) {
    k_mlp_infer_bf16<<<div_up_u(n, 64 * kWavesPerBlock), kBlock, 0, s>>>(X, n, w, Y);
}
void launch_mlp_forward_f32(hipStream_t s, const float* X, uint32_t n, const float* p, float* Y) {
    k_mlp_forward_f32<<<div_up_u(n, kBlock), kBlock, 0, s>>>(X, n, p, Y);
}
size_t mlp_scratch_floats(uint32_t B) { return (size_t)kScratchRows * B; }
uint32_t mlp_grad_chunks(uint32_t B) { return div_up_u(B, kGradChunk); }
uint32_t mlp_loss_blocks(uint32_t B) { return div_up_u(B, kBlock); }
void launch_mlp_grad(hipStream_t s, const float* X, const float* T, const uint32_t* order,
                     uint32_t B, const float* p, float* S, double* loss_part, float* partials,
                     int terms) {
    k_mlp_fwdbwd<<<div_up_u(B, kBlock), kBlock, 0, s>>>(X, T, order, B, p, S, loss_part, terms);
    dim3 g(div_up_u(kMlpParams, kBlock), mlp_grad_chunks(B));
    k_mlp_wgrad<<<g, kBlock, 0, s>>>(S, B, partials);
}
void launch_mlp_adam(hipStream_t s, float* p, const float* partials, uint32_t nchunks, float* m1,
                     float* m2, float lr, int step, float* grad_out, int apply) {
    k_mlp_adam<<<div_up_u(kMlpParams, kBlock), kBlock, 0, s>>>(p, partials, nchunks, m1, m2, lr,
                                                               0.9f, 0.999f, 1e-8f, step, grad_out,
                                                               apply);
}
void launch_mlp_adam_grad(hipStream_t s, float* p, const float* g, float* m1, float* m2, float lr,
                          int step) {
    k_mlp_adam_grad<<<div_up_u(kMlpParams, kBlock), kBlock, 0, s>>>(p, g, m1, m2, lr, 0.9f,
                                                                    0.999f, 1e-8f, step);
}
void launch_collect_pairs(hipStream_t s, const float4* pos4, const Node* nodes,
                          const BuildParams* prm, const uint32_t* sinks, uint32_t ns,
                          float inv_theta2, float inv_band2, uint32_t cap, uint32_t* pairs,
                          uint32_t* counts) {
    k_collect_pairs<<<div_up_u(ns, kBlock), kBlock, 0, s>>>(pos4, nodes, prm, sinks, ns,
                                                            inv_theta2, inv_band2, cap, pairs,
                                                            counts);
}
void launch_label_pairs(hipStream_t s, const float4* pos4, const Node* nodes, const float* feat,
                        const BuildParams* prm, const uint32_t* sinks, const uint32_t* pairs, const uint32_t* counts,
                        const uint32_t* offsets, uint32_t ns, uint32_t cap, double eps2, float* X,
                        float* T, uint32_t total) {
    if (!total) return;
    k_label_pairs<<<div_up_u(total, kWavesPerBlock), kBlock, 0, s>>>(
        pos4, nodes, feat, prm, sinks, pairs, counts, offsets, ns, cap, eps2, X, T, total);
}
void launch_cell_features(hipStream_t s, const float4* pos4, const Node* nodes, const float4* m2,
                          const void* codes, bool wide, int B, const BuildParams* prm,
                          uint32_t cap_nodes, float* feat, uint32_t* lists, uint32_t n) {
    const uint32_t wcap = cell_list_cap(n, kFeatThread), bcap = cell_list_cap(n, kFeatWave);
    uint32_t* wlist = lists;
    uint32_t* blist = lists + 1 + wcap;
    const uint32_t g = div_up_u(cap_nodes, kBlock);
    launch_fill_u32(s, wlist, 1, 0u);
    launch_fill_u32(s, blist, 1, 0u);
    if (wide)
        k_cell_features<uint64_t><<<g, kBlock, 0, s>>>(pos4, nodes, m2, (const uint64_t*)codes, B,
                                                       prm, cap_nodes, feat, wlist, wcap, blist,
                                                       bcap);
    else
        k_cell_features<uint32_t><<<g, kBlock, 0, s>>>(pos4, nodes, m2, (const uint32_t*)codes, B,
                                                       prm, cap_nodes, feat, wlist, wcap, blist,
                                                       bcap);
    k_cell_t3_wave<<<2048, kBlock, 0, s>>>(pos4, nodes, prm, wlist, wcap, feat);
    k_cell_t3_big<<<1024, kBigThreads, 0, s>>>(pos4, nodes, prm, blist, bcap, feat);
}
// cells of more than c particles are disjoint within a level: at most n / c of them per level,
// over at most kMaxLevels + 1 levels
uint32_t cell_list_cap(uint32_t n, uint32_t c) { return (n / c + 1) * (kMaxLevels + 1); }
size_t cell_features_lists_words(uint32_t n) {
    return 2 + (size_t)cell_list_cap(n, kFeatThread) + cell_list_cap(n, kFeatWave);
}
void launch_cell_coefs(hipStream_t s, const float* feat, const Node* nodes,
                       const BuildParams* prm, uint32_t cap_nodes, const MlpBf16* w,
                       float4* coef) {
    k_cell_coefs<<<div_up_u(cap_nodes, 64 * kWavesPerBlock), kBlock, 0, s>>>(feat, nodes, prm,
                                                                            cap_nodes, w, coef);
}

}  // namespace bh
