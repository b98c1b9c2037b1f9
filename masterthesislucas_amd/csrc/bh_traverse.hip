// Barnes-Hut traversal on gfx950: one wave = 64 Morton-consecutive sink particles.
//
// The wave walks the pre-order tree together: each node record is fetched once per wave with
// scalar loads, and every lane applies its own acceptance test, so each particle sees exactly
// the per-particle Barnes-Hut decisions of the reference traversal (nBodySolverBH.h:56-89 /
// nBodySolverFMM_SYM_MORTON.h:272-313). Lanes that accept a cell take its monopole and sit out
// the cell's subtree; the set of lanes still descending is kept as a 64-bit mask, and a small
// per-wave stack in LDS holds (mask, subtree end) for every level where the mask narrowed.
// Near-field cells are summed particle by particle with the self term excluded by index.
// Forces accumulate in FP32 with Kahan compensation: each particle-particle cell sum and each
// batch of at most kFoldEvery monopoles is one compensated addition.
#include "bh_common.h"
#include "bh_launch.h"

namespace bh {

__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// One interaction, accumulated into (ax, ay[, at]) for lanes where `use` holds.
// gravity: a += m (q - p) / (|q - p|^2 + eps^2)^(3/2); potential -= m / sqrt(...)
// t-SNE:   w = 1 / (1 + |p - q|^2); total += m w; derivative += m w^2 (p - q)
// (dx, dy) = q - p points from the sink to the source.
template <int KERNEL, bool TOT>
__device__ __forceinline__ void accum(float dx, float dy, float d2, float m, float eps2, bool use,
                                      float& ax, float& ay, float& at) {
    if (KERNEL == kKernelGravity) {
        const float inv = __builtin_amdgcn_rsqf(d2 + eps2);
        const float mi = m * inv;
        float s = mi * inv * inv;
        s = use ? s : 0.f;
        ax = fmaf(s, dx, ax);
        ay = fmaf(s, dy, ay);
        if (TOT) at -= use ? mi : 0.f;
    } else {
        const float w = __builtin_amdgcn_rcpf(1.0f + d2);
        const float mw = m * w;
        float s = mw * w;
        s = use ? s : 0.f;
        ax = fmaf(-s, dx, ax);
        ay = fmaf(-s, dy, ay);
        if (TOT) at += use ? mw : 0.f;
    }
}

// Wave-uniform 64-bit value moved to scalar registers (the compiler then keeps lane masks and
// node indices in SGPRs and the per-lane booleans derived from them cost no vector work).
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane(v);
}

// Traversal variants (picked per workload by the autotuner):
//   NS:      sinks per lane; a wave walks the tree once for 64 NS Morton-consecutive particles,
//            so every node fetch feeds NS independent interactions per lane
//   BRANCHY: monopoles only when some lane accepts (else computed masked, unconditionally)
//   PPU:     sources fetched per scalar load in the particle-particle loop
//
// Per-lane state that is a function of lane masks (walking / accepting / opening) is held as
// lane booleans, which live in scalar register pairs: the acceptance test is the only vector
// work per node and sink besides the interaction itself. Interaction counts are wave totals
// (popcounts of the acceptance masks).
template <int KERNEL, bool TOT, int NS, bool BRANCHY, int PPU>
__global__ __launch_bounds__(kBlock) void k_traverse(const float4* __restrict__ pos4,
                                                     const Node* __restrict__ nodes,
                                                     const BuildParams* __restrict__ prm,
                                                     uint32_t n, float inv_theta2, float eps2,
                                                     float scale, float* __restrict__ ax_out,
                                                     float* __restrict__ ay_out,
                                                     float* __restrict__ tot_out,
                                                     unsigned long long* __restrict__ counters) {
    __shared__ uint64_t st_mask[kWavesPerBlock][kStackDepth][NS];
    __shared__ uint32_t st_end[kWavesPerBlock][kStackDepth];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t wb = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t base = (wb * kWavesPerBlock + w) * (uint32_t)(kWave * NS);
    uint32_t idx[NS];
    uint64_t onm[NS];  // lanes still walking the current subtree for sink s
    float px[NS], py[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        idx[s] = base + s * kWave + lane;
        const bool valid = idx[s] < n;
        onm[s] = uniform_u64(__ballot(valid));
        const float4 me = valid ? pos4[idx[s]] : make_float4(0.f, 0.f, 0.f, 0.f);
        px[s] = me.x;
        py[s] = me.y;
    }
    const uint32_t nn = uniform_u32(prm->n_nodes);

    // Compensated totals (ax, ay, at) and their Kahan terms; monopoles of consecutive visits are
    // first gathered in a short plain batch (bx, by, bt) that is folded into the compensated
    // totals at every particle-particle visit and at least every kFoldEvery visits.
    float ax[NS], ay[NS], at[NS], cax[NS], cay[NS], cat[NS], bx[NS], by[NS], bt[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) ax[s] = ay[s] = at[s] = cax[s] = cay[s] = cat[s] = bx[s] =
        by[s] = bt[s] = 0.f;
    uint64_t cnt_pn = 0, cnt_pp = 0;  // wave totals
    uint32_t k = 0, end_cur = nn;
    int sp = 0, since_fold = 0;
    for (;;) {
        while (k >= end_cur && sp > 0) {
            --sp;
#pragma unroll
            for (int s = 0; s < NS; ++s) onm[s] = uniform_u64(st_mask[w][sp][s]);
            end_cur = uniform_u32(st_end[w][sp]);
        }
        if (k >= end_cur) break;
        uint64_t walking = 0;
#pragma unroll
        for (int s = 0; s < NS; ++s) walking |= onm[s];
        if (!walking) {  // nobody left in this subtree frame
            k = end_cur;
            continue;
        }
        const Node cur = nodes[k];
        // skip > k for every node of a well-formed tree; max() keeps the walk finite regardless
        const uint32_t next = max(cur.skip, k + 1);
        const float lim = cur.l2 * inv_theta2;
        uint64_t openm[NS];
        uint64_t open_any = 0, acc_any = 0;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const float dx = cur.cx - px[s], dy = cur.cy - py[s];
            const float d2 = fmaf(dx, dx, dy * dy);
            const uint64_t accm = __ballot(d2 > lim) & onm[s];
            if (!BRANCHY || accm) {
                const bool acc = __builtin_amdgcn_inverse_ballot_w64(accm);
                accum<KERNEL, TOT>(dx, dy, d2, cur.mass, eps2, acc, bx[s], by[s], bt[s]);
                cnt_pn += __popcll(accm);
            }
            openm[s] = onm[s] & ~accm;
            open_any |= openm[s];
            acc_any |= accm;
        }
        bool down = false;
        if (open_any) {
            if (cur.meta & kMetaPP) {
                float lx[NS], ly[NS], lt[NS];
#pragma unroll
                for (int s = 0; s < NS; ++s) lx[s] = ly[s] = lt[s] = 0.f;
                const uint32_t f = cur.first, e = f + cur.count;
                // PPU sources per scalar fetch; pos4 is padded past n, lanes mask the tail
                for (uint32_t j = f; j < e; j += PPU) {
                    float4 q[PPU];
#pragma unroll
                    for (int t = 0; t < PPU; ++t) q[t] = pos4[j + t];
#pragma unroll
                    for (int t = 0; t < PPU; ++t) {
#pragma unroll
                        for (int s = 0; s < NS; ++s) {
                            const float qx = q[t].x - px[s], qy = q[t].y - py[s];
                            const float q2 = fmaf(qx, qx, qy * qy);
                            const bool use = __builtin_amdgcn_inverse_ballot_w64(openm[s]) &&
                                             (j + t < e) && (j + t != idx[s]);
                            accum<KERNEL, TOT>(qx, qy, q2, q[t].z, eps2, use, lx[s], ly[s], lt[s]);
                        }
                    }
                }
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    const uint64_t self = __ballot(idx[s] >= f && idx[s] < e) & openm[s];
                    cnt_pp += (uint64_t)__popcll(openm[s]) * cur.count - (uint64_t)__popcll(self);
                    kahan_add(ax[s], cax[s], lx[s] + bx[s]);
                    kahan_add(ay[s], cay[s], ly[s] + by[s]);
                    if (TOT) kahan_add(at[s], cat[s], lt[s] + bt[s]);
                    bx[s] = by[s] = bt[s] = 0.f;
                }
                since_fold = 0;
            } else {
                // lanes that accepted sit out the subtree: remember the frame, narrow the masks
                if (acc_any && sp < kStackDepth) {
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        if (lane == 0) st_mask[w][sp][s] = onm[s];
                        onm[s] = openm[s];
                    }
                    if (lane == 0) st_end[w][sp] = end_cur;
                    ++sp;
                    end_cur = next;
                }
                down = true;
            }
        }
        if (++since_fold >= kFoldEvery) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                kahan_add(ax[s], cax[s], bx[s]);
                kahan_add(ay[s], cay[s], by[s]);
                if (TOT) kahan_add(at[s], cat[s], bt[s]);
                bx[s] = by[s] = bt[s] = 0.f;
            }
            since_fold = 0;
        }
        k = down ? k + 1 : next;
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        kahan_add(ax[s], cax[s], bx[s]);
        kahan_add(ay[s], cay[s], by[s]);
        if (TOT) kahan_add(at[s], cat[s], bt[s]);
        if (idx[s] < n) {
            ax_out[idx[s]] = scale * (ax[s] - cax[s]);
            ay_out[idx[s]] = scale * (ay[s] - cay[s]);
            if (TOT) tot_out[idx[s]] = scale * (at[s] - cat[s]);
        }
    }
    if (counters) {  // one partial per block (no same-address device atomics)
        __shared__ unsigned long long s_cnt[kWavesPerBlock][2];
        if (lane == 0) {
            s_cnt[w][0] = cnt_pn;
            s_cnt[w][1] = cnt_pp;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long a = 0, b = 0;
            for (int k2 = 0; k2 < kWavesPerBlock; ++k2) {
                a += s_cnt[k2][0];
                b += s_cnt[k2][1];
            }
            counters[2 * (size_t)wb] = a;
            counters[2 * (size_t)wb + 1] = b;
        }
    }
}

// ------------------------------------------------------------------------------------------
// FP64 reduction of per-particle totals (t-SNE Z, potential energy): fixed partition into
// block partials, then one block sums them in order, so the result is deterministic.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ double block_sum_d(double s) {
    __shared__ double red[kWavesPerBlock];
    s = wave_sum_d(s);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = s;
    __syncthreads();
    double b = 0.0;
    for (int k = 0; k < kWavesPerBlock; ++k) b += red[k];
    __syncthreads();
    return b;
}

__global__ __launch_bounds__(kBlock) void k_reduce_f64_partial(const float* __restrict__ v,
                                                               uint32_t n,
                                                               const float* __restrict__ w,
                                                               double* __restrict__ partials) {
    double s = 0.0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        s += w ? (double)v[i] * (double)w[i] : (double)v[i];
    s = block_sum_d(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_reduce_f64_final(const double* __restrict__ partials,
                                                             uint32_t nb, double* __restrict__ out) {
    double s = 0.0;
    for (uint32_t b = threadIdx.x; b < nb; b += kBlock) s += partials[b];
    s = block_sum_d(s);
    if (threadIdx.x == 0) *out = s;
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

// variant v: sinks per lane {1, 2, 4}[v / 4]; monopoles branchy (v % 4 < 2) or masked;
// particle-particle fetch of 1 (v even) or 4 (v odd) sources
template <int KERNEL, bool TOT>
static uint32_t traverse_variant(hipStream_t s, int variant, uint32_t n, const float4* pos4,
                                 const Node* nodes, const BuildParams* prm, float inv_theta2,
                                 float eps2, float scale, float* ax, float* ay, float* tot,
                                 unsigned long long* counters) {
    const int ns = variant < 4 ? 1 : (variant < 8 ? 2 : 4);
    const uint32_t g = div_up(n, (uint64_t)kBlock * ns);
#define BH_TRAV(NS, BR, U)                                                                        \
    k_traverse<KERNEL, TOT, NS, BR, U><<<g, kBlock, 0, s>>>(pos4, nodes, prm, n, inv_theta2,     \
                                                             eps2, scale, ax, ay, tot, counters)
    switch (variant) {
        case 0: BH_TRAV(1, true, 1); break;
        case 1: BH_TRAV(1, true, 4); break;
        case 2: BH_TRAV(1, false, 1); break;
        case 3: BH_TRAV(1, false, 4); break;
        case 4: BH_TRAV(2, true, 1); break;
        case 5: BH_TRAV(2, true, 4); break;
        case 6: BH_TRAV(2, false, 1); break;
        case 7: BH_TRAV(2, false, 4); break;
        case 8: BH_TRAV(4, true, 1); break;
        case 9: BH_TRAV(4, true, 4); break;
        case 10: BH_TRAV(4, false, 1); break;
        default: BH_TRAV(4, false, 4); break;
    }
#undef BH_TRAV
    return g;
}

uint32_t launch_traverse(hipStream_t s, int kernel, int variant, const float4* pos4,
                         const Node* nodes, const BuildParams* prm, uint32_t n, float inv_theta2,
                         float eps2, float scale, float* ax, float* ay, float* tot,
                         unsigned long long* counters) {
    if (kernel == kKernelGravity) {
        if (tot)
            return traverse_variant<kKernelGravity, true>(s, variant, n, pos4, nodes, prm,
                                                          inv_theta2, eps2, scale, ax, ay, tot,
                                                          counters);
        return traverse_variant<kKernelGravity, false>(s, variant, n, pos4, nodes, prm, inv_theta2,
                                                       eps2, scale, ax, ay, tot, counters);
    }
    if (tot)
        return traverse_variant<kKernelTsne, true>(s, variant, n, pos4, nodes, prm, inv_theta2,
                                                   eps2, scale, ax, ay, tot, counters);
    return traverse_variant<kKernelTsne, false>(s, variant, n, pos4, nodes, prm, inv_theta2, eps2,
                                                scale, ax, ay, tot, counters);
}

uint32_t reduce_grid(uint32_t n) {
    uint32_t g = div_up(n, kBlock * 16);
    if (g < 1) g = 1;
    if (g > kMaxReduceBlocks) g = kMaxReduceBlocks;
    return g;
}

void launch_reduce_f64(hipStream_t s, const float* v, const float* w, uint32_t n, double* partials,
                       double* out) {
    const uint32_t g = reduce_grid(n);
    k_reduce_f64_partial<<<g, kBlock, 0, s>>>(v, n, w, partials);
    k_reduce_f64_final<<<1, kBlock, 0, s>>>(partials, g, out);
}

uint32_t traverse_blocks(uint32_t n) { return div_up(n, kBlock); }

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
