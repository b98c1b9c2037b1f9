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
        for (int k = 0; k < 4; ++k) {
            bf16x4 v;
            for (int i = 0; i < 4; ++i) v[i] = f32_to_bf16(W2[(16 * k + 4 * q4 + i) * kMlpH + 16 * u + r16]);
            W->w2[u][k][lane] = v;
        }
    for (int k = 0; k < 4; ++k) {
        bf16x4 v;
        for (int i = 0; i < 4; ++i)
            v[i] = r16 < kMlpOut ? f32_to_bf16(W3[(16 * k + 4 * q4 + i) * kMlpOut + r16]) : (short)0;
        W->w3[k][lane] = v;
    }
    W->b2[lane] = p[kMlpOffB2 + lane];
    if (lane < kMlpOut) W->b3[lane] = p[kMlpOffB3 + lane];
}

// ---------------------------------------------------------------------------------------------
// inference: BF16 MFMA (one wave per 64 samples) and the FP32 reference forward
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_mlp_infer_bf16(const float* __restrict__ X, uint32_t n,
                                                           const MlpBf16* __restrict__ W,
                                                           float* __restrict__ Y) {
    __shared__ short xs[kWavesPerBlock][64 * kMlpInPad];
    __shared__ MlpBf16 ws;
    // stage the weights once per block
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(W);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&ws);
        for (uint32_t i = threadIdx.x; i < sizeof(MlpBf16) / 4; i += kBlock) dst[i] = src[i];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t base = (blockIdx.x * kWavesPerBlock + w) * 64u;
    const uint32_t i = base + lane;
    short* row = &xs[w][lane * kMlpInPad];
#pragma unroll
    for (int c = 0; c < kMlpInPad; ++c) {
        float v = 0.f;
        if (c < kMlpIn) v = i < n ? X[(size_t)i * kMlpIn + c] : 0.f;
        else if (c == kMlpIn) v = 1.f;
        row[c] = f32_to_bf16(v);
    }
    __syncthreads();
    float o0[4], o1[4];
    mlp_wave_infer(&ws, xs[w], o0, o1);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const uint32_t j = base + 16 * s + lane;
        if (lane < 16 && j < n) {
            Y[(size_t)j * 2] = o0[s];
            Y[(size_t)j * 2 + 1] = o1[s];
        }
    }
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
    float x[kMlpIn], a1[kMlpH];
#pragma unroll
    for (int c = 0; c < kMlpIn; ++c) x[c] = X[(size_t)i * kMlpIn + c];
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) {
        float z = p[kMlpOffB1 + j];
#pragma unroll
        for (int c = 0; c < kMlpIn; ++c) z = fmaf(x[c], p[kMlpOffW1 + c * kMlpH + j], z);
        a1[j] = silu_f(z);
    }
    float y0 = p[kMlpOffB3], y1 = p[kMlpOffB3 + 1];
    for (int j = 0; j < kMlpH; ++j) {
        float z = p[kMlpOffB2 + j];
#pragma unroll
        for (int k = 0; k < kMlpH; ++k) z = fmaf(a1[k], p[kMlpOffW2 + k * kMlpH + j], z);
        const float a2 = silu_f(z);
        y0 = fmaf(a2, p[kMlpOffW3 + j * 2], y0);
        y1 = fmaf(a2, p[kMlpOffW3 + j * 2 + 1], y1);
    }
    Y[(size_t)i * 2] = y0;
    Y[(size_t)i * 2 + 1] = y1;
}

// ---------------------------------------------------------------------------------------------
// training: per-sample forward + backward to the pre-activation gradients, written
// feature-major to scratch S (rows of length B): A1[64] DZ1[64] A2[64] DZ2[64] DY[2] X[10]
// ---------------------------------------------------------------------------------------------
constexpr int kRowA1 = 0, kRowDZ1 = 64, kRowA2 = 128, kRowDZ2 = 192, kRowDY = 256, kRowX = 258;
constexpr int kScratchRows = 268;

__global__ __launch_bounds__(kBlock) void k_mlp_fwdbwd(const float* __restrict__ X,
                                                       const float* __restrict__ T,
                                                       const uint32_t* __restrict__ order,
                                                       uint32_t B, const float* __restrict__ p,
                                                       float* __restrict__ S,
                                                       double* __restrict__ loss_part) {
    __shared__ double red[kWavesPerBlock];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    double sq = 0.0;
    if (i < B) {
        const uint32_t src = order ? order[i] : i;
        float x[kMlpIn];
#pragma unroll
        for (int c = 0; c < kMlpIn; ++c) {
            x[c] = X[(size_t)src * kMlpIn + c];
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
        float y0 = p[kMlpOffB3], y1 = p[kMlpOffB3 + 1];
        for (int j = 0; j < kMlpH; ++j) {
            float z = p[kMlpOffB2 + j];
#pragma unroll
            for (int k = 0; k < kMlpH; ++k) z = fmaf(a1[k], p[kMlpOffW2 + k * kMlpH + j], z);
            const float a2 = silu_f(z);
            S[(size_t)(kRowA2 + j) * B + i] = a2;
            S[(size_t)(kRowDZ2 + j) * B + i] = dsilu_f(z);  // silu'(z2), scaled below
            y0 = fmaf(a2, p[kMlpOffW3 + j * 2], y0);
            y1 = fmaf(a2, p[kMlpOffW3 + j * 2 + 1], y1);
        }
        const float e0 = y0 - T[(size_t)src * 2], e1 = y1 - T[(size_t)src * 2 + 1];
        sq = (double)e0 * e0 + (double)e1 * e1;
        // loss = sum(e^2) / (2 B): dL/dy = e / B
        const float dy0 = e0 / (float)B, dy1 = e1 / (float)B;
        S[(size_t)kRowDY * B + i] = dy0;
        S[(size_t)(kRowDY + 1) * B + i] = dy1;
        float dz2[kMlpH];
#pragma unroll
        for (int j = 0; j < kMlpH; ++j) {
            const float da2 = fmaf(dy0, p[kMlpOffW3 + j * 2], dy1 * p[kMlpOffW3 + j * 2 + 1]);
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
// index) minus the FP64 monopole, features in FP32.
__global__ __launch_bounds__(kBlock) void k_label_pairs(const float4* __restrict__ pos4,
                                                        const Node* __restrict__ nodes,
                                                        const float4* __restrict__ m2,
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
    const double rx = sx - sm * dx, ry = sy - sm * dy;
    float f[kMlpIn];
    const float4 mm = m2[k];
    const float fdx = nd.cx - me.x, fdy = nd.cy - me.y;
    const float scale = mlp_features(fdx, fdy, fmaf(fdx, fdx, fdy * fdy), nd.l2, nd.mass, nd.count,
                                     mm.x, mm.y, mm.z, f);
#pragma unroll
    for (int q = 0; q < kMlpIn; ++q) Xo[(size_t)gw * kMlpIn + q] = f[q];
    To[(size_t)gw * 2] = (float)(rx / (double)scale);
    To[(size_t)gw * 2 + 1] = (float)(ry / (double)scale);
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
                     uint32_t B, const float* p, float* S, double* loss_part, float* partials) {
    k_mlp_fwdbwd<<<div_up_u(B, kBlock), kBlock, 0, s>>>(X, T, order, B, p, S, loss_part);
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
void launch_label_pairs(hipStream_t s, const float4* pos4, const Node* nodes, const float4* m2,
                        const uint32_t* sinks, const uint32_t* pairs, const uint32_t* counts,
                        const uint32_t* offsets, uint32_t ns, uint32_t cap, double eps2, float* X,
                        float* T, uint32_t total) {
    if (!total) return;
    k_label_pairs<<<div_up_u(total, kWavesPerBlock), kBlock, 0, s>>>(
        pos4, nodes, m2, sinks, pairs, counts, offsets, ns, cap, eps2, X, T, total);
}

}  // namespace bh
