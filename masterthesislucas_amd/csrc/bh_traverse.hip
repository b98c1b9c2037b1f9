// Barnes-Hut traversal on gfx950: one wave = 64 Morton-consecutive sink particles.
//
// The wave walks the pre-order tree together: each node record is fetched once per wave with
// scalar loads, and every lane applies its own acceptance test, so each particle sees exactly
// the per-particle Barnes-Hut decisions of the reference traversal (nBodySolverBH.h:56-89 /
// nBodySolverFMM_SYM_MORTON.h:272-313). Lanes that accept a cell take its monopole and sit out
// the cell's subtree; the set of lanes still descending is kept as a 64-bit mask, and a small
// per-wave stack in LDS holds (mask, subtree end) for every level where the mask narrowed.
// Near-field cells are summed particle by particle with the self term excluded by index.
// Forces accumulate in FP32 with Kahan compensation per node visit.
#include "bh_common.h"
#include "bh_launch.h"

namespace bh {

__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// gravity: a += m (q - p) / (|q - p|^2 + eps^2)^(3/2); potential -= m / sqrt(...)
// t-SNE:   w = 1 / (1 + |p - q|^2); total += m w; derivative += m w^2 (p - q)
template <int KERNEL>
__device__ __forceinline__ void pair_term(float dx, float dy, float d2, float m, float eps2,
                                          float& tx, float& ty, float& tt) {
    if (KERNEL == kKernelGravity) {
        const float inv = __builtin_amdgcn_rsqf(d2 + eps2);
        const float mi = m * inv;
        const float s = mi * inv * inv;
        tx = s * dx;
        ty = s * dy;
        tt = -mi;
    } else {
        const float w = __builtin_amdgcn_rcpf(1.0f + d2);
        const float mw = m * w;
        const float s = mw * w;
        tx = -s * dx;  // dx, dy point from the sink to the source; the t-SNE term uses p - q
        ty = -s * dy;
        tt = mw;
    }
}

template <int KERNEL>
__global__ __launch_bounds__(kBlock) void k_traverse(const float4* __restrict__ pos4,
                                                     const Node* __restrict__ nodes,
                                                     const BuildParams* __restrict__ prm,
                                                     uint32_t n, float inv_theta2, float eps2,
                                                     float scale, float* __restrict__ ax_out,
                                                     float* __restrict__ ay_out,
                                                     float* __restrict__ tot_out,
                                                     unsigned long long* __restrict__ counters) {
    __shared__ uint64_t st_mask[kWavesPerBlock][kStackDepth];
    __shared__ uint32_t st_end[kWavesPerBlock][kStackDepth];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t wb = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t i = (wb * kWavesPerBlock + w) * kWave + lane;
    const bool valid = i < n;
    const float4 me = valid ? pos4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    uint64_t active = __ballot(valid);
    if (!active) return;
    const uint32_t nn = prm->n_nodes;

    float ax = 0.f, ay = 0.f, at = 0.f, cax = 0.f, cay = 0.f, cat = 0.f;
    uint32_t npn = 0, npp = 0;
    uint32_t k = 0, end_cur = nn;
    int sp = 0;
    while (true) {
        while (k == end_cur && sp > 0) {
            --sp;
            active = st_mask[w][sp];
            end_cur = st_end[w][sp];
        }
        if (k >= end_cur) break;
        const Node nd = nodes[k];
        const bool on = (active >> lane) & 1ull;
        const float dx = nd.cx - me.x, dy = nd.cy - me.y;
        const float d2 = fmaf(dx, dx, dy * dy);
        const bool acc = on && (d2 > nd.l2 * inv_theta2);
        if (acc) {
            float tx, ty, tt;
            pair_term<KERNEL>(dx, dy, d2, nd.mass, eps2, tx, ty, tt);
            kahan_add(ax, cax, tx);
            kahan_add(ay, cay, ty);
            kahan_add(at, cat, tt);
            ++npn;
        }
        const uint64_t open = __ballot(on && !acc);
        // skip > k for every node of a well-formed tree; max() keeps the walk finite regardless
        const uint32_t next = max(nd.skip, k + 1);
        if (!open) {
            k = next;
            continue;
        }
        if (nd.meta & kMetaPP) {
            if ((open >> lane) & 1ull) {
                float lx = 0.f, ly = 0.f, lt = 0.f;
                const uint32_t f = nd.first, e = f + nd.count;
                for (uint32_t j = f; j < e; ++j) {
                    const float4 q = pos4[j];
                    const float qx = q.x - me.x, qy = q.y - me.y;
                    const float q2 = fmaf(qx, qx, qy * qy);
                    float tx, ty, tt;
                    pair_term<KERNEL>(qx, qy, q2, q.z, eps2, tx, ty, tt);
                    if (j != i) {
                        lx += tx;
                        ly += ty;
                        lt += tt;
                    }
                }
                kahan_add(ax, cax, lx);
                kahan_add(ay, cay, ly);
                kahan_add(at, cat, lt);
                npp += nd.count - ((i >= f && i < e) ? 1u : 0u);
            }
            k = next;
            continue;
        }
        if (open != active && sp < kStackDepth) {
            if (lane == 0) {
                st_mask[w][sp] = active;
                st_end[w][sp] = end_cur;
            }
            ++sp;
            active = open;
            end_cur = next;
        }
        k = k + 1;
    }
    if (valid) {
        ax_out[i] = scale * (ax - cax);
        ay_out[i] = scale * (ay - cay);
        if (tot_out) tot_out[i] = scale * (at - cat);
    }
    if (counters) {
        const unsigned long long s_pn = wave_sum_u64(valid ? npn : 0u);
        const unsigned long long s_pp = wave_sum_u64(valid ? npp : 0u);
        if (lane == 0) {
            atomicAdd(&counters[0], s_pn);
            atomicAdd(&counters[1], s_pp);
        }
    }
}

// ------------------------------------------------------------------------------------------
// FP64 reduction of per-particle totals (t-SNE Z, potential energy): fixed partition, the
// last block sums the block partials in order, so the result is deterministic.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_reduce_f64(const float* __restrict__ v, uint32_t n,
                                                       const float* __restrict__ w,
                                                       double* __restrict__ partials,
                                                       uint32_t* __restrict__ counter,
                                                       double* __restrict__ out) {
    __shared__ double red[kWavesPerBlock];
    __shared__ int last;
    double s = 0.0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        s += w ? (double)v[i] * (double)w[i] : (double)v[i];
    s = wave_sum_d(s);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = 0.0;
        for (int k = 0; k < kWavesPerBlock; ++k) b += red[k];
        partials[blockIdx.x] = b;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        last = (atomicAdd(counter, 1u) == gridDim.x - 1);
    }
    __syncthreads();
    if (!last || threadIdx.x != 0) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    double t = 0.0;
    for (uint32_t b = 0; b < gridDim.x; ++b)
        t += __hip_atomic_load(&partials[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *out = t;
    *counter = 0u;
}

// out[id[i]] = a[i] (results back to the caller's particle order)
__global__ __launch_bounds__(kBlock) void k_unsort2(const uint32_t* __restrict__ id, uint32_t n,
                                                    const float* __restrict__ a,
                                                    const float* __restrict__ b,
                                                    float* __restrict__ ao, float* __restrict__ bo) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = id[i];
    ao[j] = a[i];
    if (b) bo[j] = b[i];
}

// kick-drift of kick-drift-kick leapfrog: v += a dt/2; x += v dt
__global__ __launch_bounds__(kBlock) void k_kick_drift(float* __restrict__ x, float* __restrict__ y,
                                                       float* __restrict__ vx,
                                                       float* __restrict__ vy,
                                                       const float* __restrict__ ax,
                                                       const float* __restrict__ ay, uint32_t n,
                                                       float hdt, float dt) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float nvx = fmaf(ax[i], hdt, vx[i]);
    const float nvy = fmaf(ay[i], hdt, vy[i]);
    vx[i] = nvx;
    vy[i] = nvy;
    x[i] = fmaf(nvx, dt, x[i]);
    y[i] = fmaf(nvy, dt, y[i]);
}

__global__ __launch_bounds__(kBlock) void k_kick(float* __restrict__ vx, float* __restrict__ vy,
                                                 const float* __restrict__ ax,
                                                 const float* __restrict__ ay, uint32_t n,
                                                 float hdt) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    vx[i] = fmaf(ax[i], hdt, vx[i]);
    vy[i] = fmaf(ay[i], hdt, vy[i]);
}

// ------------------------------------------------------------------------------------------
static inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

void launch_traverse(hipStream_t s, int kernel, const float4* pos4, const Node* nodes,
                     const BuildParams* prm, uint32_t n, float inv_theta2, float eps2, float scale,
                     float* ax, float* ay, float* tot, unsigned long long* counters) {
    const uint32_t g = div_up(n, kBlock);
    if (kernel == kKernelGravity)
        k_traverse<kKernelGravity><<<g, kBlock, 0, s>>>(pos4, nodes, prm, n, inv_theta2, eps2, scale,
                                                        ax, ay, tot, counters);
    else
        k_traverse<kKernelTsne><<<g, kBlock, 0, s>>>(pos4, nodes, prm, n, inv_theta2, eps2, scale,
                                                     ax, ay, tot, counters);
}

uint32_t reduce_grid(uint32_t n) {
    uint32_t g = div_up(n, kBlock * 16);
    if (g < 1) g = 1;
    if (g > kMaxReduceBlocks) g = kMaxReduceBlocks;
    return g;
}

void launch_reduce_f64(hipStream_t s, const float* v, const float* w, uint32_t n, double* partials,
                       uint32_t* counter, double* out) {
    k_reduce_f64<<<reduce_grid(n), kBlock, 0, s>>>(v, n, w, partials, counter, out);
}

void launch_unsort2(hipStream_t s, const uint32_t* id, uint32_t n, const float* a, const float* b,
                    float* ao, float* bo) {
    k_unsort2<<<div_up(n, kBlock), kBlock, 0, s>>>(id, n, a, b, ao, bo);
}

void launch_kick_drift(hipStream_t s, float* x, float* y, float* vx, float* vy, const float* ax,
                       const float* ay, uint32_t n, float hdt, float dt) {
    k_kick_drift<<<div_up(n, kBlock), kBlock, 0, s>>>(x, y, vx, vy, ax, ay, n, hdt, dt);
}

void launch_kick(hipStream_t s, float* vx, float* vy, const float* ax, const float* ay, uint32_t n,
                 float hdt) {
    k_kick<<<div_up(n, kBlock), kBlock, 0, s>>>(vx, vy, ax, ay, n, hdt);
}

}  // namespace bh
