// t-SNE gradient step on gfx950 (the TSNE driver of the reference, tsne.h:179-233, 272-301,
// 356-460), built on the Barnes-Hut repulsion of this library.
//
// The embedding lives in the caller's point order (P is indexed that way). Per iteration:
// stage the embedding as the solver's particles, build the tree, traverse with the t-SNE kernel
// (repulsion and per-point Z terms in sorted order), reduce Z in FP64, bring the repulsion
// back to the caller's order, then one kernel per row adds the CSR attraction and applies the
// momentum update y' = y + lr d + mu (y - y_prev).
#include "bh_common.h"
#include "bh_launch.h"

namespace bh {

// embedding -> solver particles (unit masses, identity ids)
__global__ __launch_bounds__(kBlock) void k_tsne_stage(const float* __restrict__ yx,
                                                       const float* __restrict__ yy, uint32_t n,
                                                       float* __restrict__ x, float* __restrict__ y,
                                                       float* __restrict__ m,
                                                       uint32_t* __restrict__ id,
                                                       float2* __restrict__ xy) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float a = yx[i], b = yy[i];
    x[i] = a;
    y[i] = b;
    xy[i] = make_float2(a, b);
    m[i] = 1.0f;
    id[i] = i;
}

// One wave per row i: attraction exag * sum_j P_ij w_ij (y_j - y_i) with w = 1/(1+|y_j-y_i|^2)
// (updateAttractive, tsne.h:356-381). It needs only the current embedding, so it runs on a
// stream of its own while the tree is built and walked. The neighbours' positions are gathered
// from the interleaved copy xy (one 8-byte load per non-zero instead of two 4-byte ones: the
// gathers, not the bytes, bound this kernel).
__global__ __launch_bounds__(kBlock) void k_tsne_attract(const uint32_t* __restrict__ row_ptr,
                                                         const uint32_t* __restrict__ col,
                                                         const float* __restrict__ val, uint32_t n,
                                                         const float2* __restrict__ xy, float exag,
                                                         float* __restrict__ att_x,
                                                         float* __restrict__ att_y) {
    const int lane = threadIdx.x & 63;
    const uint32_t r = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (r >= n) return;
    const float2 yi = xy[r];
    float sx = 0.f, sy = 0.f;
    const uint32_t e0 = row_ptr[r], e1 = row_ptr[r + 1];
    for (uint32_t e = e0 + lane; e < e1; e += kWave) {
        const float2 yc = xy[col[e]];
        const float ddx = yc.x - yi.x, ddy = yc.y - yi.y;
        const float q = __builtin_amdgcn_rcpf(1.0f + fmaf(ddx, ddx, ddy * ddy));
        const float s = exag * val[e] * q;
        sx = fmaf(s, ddx, sx);
        sy = fmaf(s, ddy, sy);
    }
    sx = wave_sum_f(sx);
    sy = wave_sum_f(sy);
    if (lane != 0) return;
    att_x[r] = sx;
    att_y[r] = sy;
}

// derivative d_i = (s/Z) rep_i + attraction with s = 4 (updateRepulsive scaling,
// tsne.h:291-294; 1 in the reference's GPU step, nBodySolverGpuBH.cu:424), then
// y' = y + lr d + mu (y - y_prev) (updatePoints, tsne.h:179-206); one thread per point
__global__ __launch_bounds__(kBlock) void k_tsne_combine(uint32_t n, const float* __restrict__ yx,
                                                         const float* __restrict__ yy,
                                                         const float* __restrict__ pyx,
                                                         const float* __restrict__ pyy,
                                                         const float* __restrict__ rep_x,
                                                         const float* __restrict__ rep_y,
                                                         const float* __restrict__ att_x,
                                                         const float* __restrict__ att_y,
                                                         const double* __restrict__ z,
                                                         double rep_scale, float lr, float mu,
                                                         float* __restrict__ nyx,
                                                         float* __restrict__ nyy,
                                                         float* __restrict__ dx_out,
                                                         float* __restrict__ dy_out) {
    const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    const float scale = (float)(rep_scale / *z);
    const float dx = fmaf(scale, rep_x[r], att_x[r]);
    const float dy = fmaf(scale, rep_y[r], att_y[r]);
    if (dx_out) {
        dx_out[r] = dx;
        dy_out[r] = dy;
    }
    if (nyx) {
        const float xi = yx[r], yi = yy[r];
        nyx[r] = fmaf(lr, dx, xi) + mu * (xi - pyx[r]);
        nyy[r] = fmaf(lr, dy, yi) + mu * (yi - pyy[r]);
    }
}

// KL divergence sum_{i != j} P_ij log2(P_ij / Q_ij), Q_ij = w_ij / Z, in FP64
// (costFunction, tsne.h:383-460); block partials, summed in order by k_tsne_kl_final.
__global__ __launch_bounds__(kBlock) void k_tsne_kl(const uint32_t* __restrict__ row_ptr,
                                                    const uint32_t* __restrict__ col,
                                                    const float* __restrict__ val, uint32_t n,
                                                    const float* __restrict__ yx,
                                                    const float* __restrict__ yy,
                                                    const double* __restrict__ z,
                                                    double* __restrict__ partials) {
    __shared__ double red[kWavesPerBlock];
    const double Z = *z;
    double acc = 0.0;
    for (uint32_t r = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); r < n;
         r += gridDim.x * kWavesPerBlock) {
        const double xi = yx[r], yi = yy[r];
        for (uint32_t e = row_ptr[r] + (threadIdx.x & 63); e < row_ptr[r + 1]; e += kWave) {
            const uint32_t c = col[e];
            if (c == r) continue;
            const double ddx = (double)yx[c] - xi, ddy = (double)yy[c] - yi;
            const double q = (1.0 / (1.0 + (ddx * ddx + ddy * ddy))) / Z;
            const double p = val[e];
            acc += p * log2(p / q);
        }
    }
    acc = wave_sum_d(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = 0.0;
        for (int k = 0; k < kWavesPerBlock; ++k) b += red[k];
        partials[blockIdx.x] = b;
    }
}

__global__ __launch_bounds__(kBlock) void k_sum_f64(const double* __restrict__ v, uint32_t nv,
                                                    double* __restrict__ out) {
    __shared__ double red[kWavesPerBlock];
    double s = 0.0;
    for (uint32_t b = threadIdx.x; b < nv; b += kBlock) s += v[b];
    s = wave_sum_d(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = 0.0;
        for (int k = 0; k < kWavesPerBlock; ++k) b += red[k];
        *out = b;
    }
}

static inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

void launch_tsne_stage(hipStream_t s, const float* yx, const float* yy, uint32_t n, float* x,
                       float* y, float* m, uint32_t* id, float2* xy) {
    k_tsne_stage<<<div_up(n, kBlock), kBlock, 0, s>>>(yx, yy, n, x, y, m, id, xy);
}

void launch_tsne_attract(hipStream_t s, const uint32_t* row_ptr, const uint32_t* col,
                         const float* val, uint32_t n, const float2* xy, float exag, float* att_x,
                         float* att_y) {
    k_tsne_attract<<<div_up(n, kWavesPerBlock), kBlock, 0, s>>>(row_ptr, col, val, n, xy, exag,
                                                                att_x, att_y);
}

void launch_tsne_combine(hipStream_t s, uint32_t n, const float* yx, const float* yy,
                         const float* pyx, const float* pyy, const float* rep_x,
                         const float* rep_y, const float* att_x, const float* att_y,
                         const double* z, double rep_scale, float lr, float mu, float* nyx,
                         float* nyy, float* dx, float* dy) {
    k_tsne_combine<<<div_up(n, kBlock), kBlock, 0, s>>>(n, yx, yy, pyx, pyy, rep_x, rep_y, att_x,
                                                        att_y, z, rep_scale, lr, mu, nyx, nyy, dx,
                                                        dy);
}

uint32_t tsne_kl_blocks(uint32_t n) {
    uint32_t g = div_up(n, kWavesPerBlock);
    return g > kMaxReduceBlocks ? kMaxReduceBlocks : (g ? g : 1);
}

void launch_tsne_kl(hipStream_t s, const uint32_t* row_ptr, const uint32_t* col, const float* val,
                    uint32_t n, const float* yx, const float* yy, const double* z,
                    double* partials, double* out) {
    const uint32_t g = tsne_kl_blocks(n);
    k_tsne_kl<<<g, kBlock, 0, s>>>(row_ptr, col, val, n, yx, yy, z, partials);
    k_sum_f64<<<1, kBlock, 0, s>>>(partials, g, out);
}

void launch_sum_f64(hipStream_t s, const double* v, uint32_t nv, double* out) {
    k_sum_f64<<<1, kBlock, 0, s>>>(v, nv, out);
}

}  // namespace bh
