// t-SNE input affinities on gfx950: exact k nearest neighbours, a Gaussian bandwidth per point
// found by bisection to a perplexity, and the symmetrised sparse P = (P_cond + P_cond^T) / (2n)
// in CSR sorted by (row, column).
//
// The reference prepares P offline (MnistLoader.ipynb:539-547, openTSNE's exact
// PerplexityBasedNN and joint_probabilities_nn) and loads it from a Matrix Market file
// (tsne.h:94-98); setupEfficientPmatrix keeps it as row-sorted CSR (tsne.h:462-520).
//
// Neighbours. Candidates come from FP16 Gram products on MFMA (v_mfma_f32_16x16x32_f16, FP32
// accumulation), exact ones from FP64 distances of the FP32 inputs. Each (i, j) gets rigorous
// bounds [lo, hi] of the exact Euclidean distance: with x^ the FP16 rounding of the scaled
// point x (|x - x^| <= 2^-11 |x| + sqrt(dim) 2^-25, scaled units, RNE with subnormals) and d^
// the FP32 value of |x^_i|^2 + |x^_j|^2 - 2 x^_i.x^_j (error at most
// delta = 2^-24 ((2K + 4) |x^_i| |x^_j| + 4 (|x^_i|^2 + |x^_j|^2)) for K accumulated products),
// |x_i - x_j| lies in [sqrt(d^ - delta) - e_i - e_j, sqrt(d^ + delta) + e_i + e_j]. A row keeps
// every column whose lower bound does not exceed T, the k-th smallest upper bound seen so far,
// so the final candidates contain the exact k nearest; their FP64 distances decide, ties by
// index. A row whose candidate window outgrows its buffer is reported (an error, never a
// silently wrong neighbour list).
#include "bh_common.h"
#include "bh_launch.h"

namespace bh {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ static inline uint32_t div_up_a(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

constexpr int kAffCols = 64;                  // candidate columns per tile
constexpr int kAffKC = 128;                   // K halfs per staged chunk
constexpr int kAffPad = 8;                    // LDS row padding (conflict-free 16-byte reads)
constexpr uint32_t kAffCap = 256;             // candidate slots per row
constexpr uint32_t kAffKeep = kAffCap - kAffCols;  // compact above this many (a tile adds <= 64)
constexpr int kAffMaxDimPad = 1024;           // dim <= 1024 (8 chunks of operands in registers)

// ---------------------------------------------------------------------------------------------
// scaled FP16 copy, squared norms and rounding bounds (one wave per point)
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_aff_pack(const float* __restrict__ X, uint32_t n,
                                                     uint32_t dim, uint32_t dp, float scale,
                                                     _Float16* __restrict__ Xh,
                                                     float* __restrict__ nh,
                                                     float* __restrict__ ah,
                                                     float* __restrict__ eb) {
    const int lane = threadIdx.x & 63;
    const uint32_t i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (i >= n) return;
    double sh = 0.0, sx = 0.0;
    for (uint32_t e = lane; e < dp; e += kWave) {
        const float v = e < dim ? X[(size_t)i * dim + e] * scale : 0.f;  // scale = 2^m
        const _Float16 h = (_Float16)v;
        Xh[(size_t)i * dp + e] = h;
        const double hd = (double)h;
        sh += hd * hd;
        sx += (double)v * (double)v;
    }
    sh = wave_sum_d(sh);
    sx = wave_sum_d(sx);
    if (lane == 0) {
        nh[i] = (float)sh;
        ah[i] = (float)sqrt(sh) * (1.0f + 0x1p-20f);
        eb[i] = (float)(0x1p-11 * sqrt(sx) + sqrt((double)dim) * 0x1p-25) * (1.0f + 0x1p-20f);
    }
}

// ---------------------------------------------------------------------------------------------
// candidates: a workgroup owns 64 query rows (their FP16 block resident in LDS) and streams all
// columns in tiles of 64; per tile each wave computes a 16 x 64 block of Gram products
// ---------------------------------------------------------------------------------------------
// compact row r's candidate buffer (one wave): T = min(T, k-th smallest upper bound), keep the
// entries whose lower bound is at most T, in order
__device__ __forceinline__ void aff_compact(uint32_t* __restrict__ cj, float* __restrict__ clo,
                                            float* __restrict__ chi, uint32_t* cnt, float* T,
                                            uint32_t* over, uint32_t k) {
    const int lane = threadIdx.x & 63;
    const uint32_t c = *cnt;
    uint32_t j[4];
    float lo[4];
    uint32_t hb[4];
    bool v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const uint32_t q = lane + 64u * u;
        v[u] = q < c;
        j[u] = v[u] ? cj[q] : 0u;
        lo[u] = v[u] ? clo[q] : 0.f;
        hb[u] = v[u] ? __float_as_uint(chi[q]) : 0u;  // upper bounds are >= 0: bits order them
    }
    float t = *T;
    if (c >= k) {
        uint32_t a = 0u, b = 0x7F800000u;  // smallest pattern with #(hi <= pattern) >= k
        while (a < b) {
            const uint32_t mid = a + ((b - a) >> 1);
            uint32_t m = 0;
#pragma unroll
            for (int u = 0; u < 4; ++u) m += (uint32_t)__popcll(__ballot(v[u] && hb[u] <= mid));
            if (m >= k) b = mid; else a = mid + 1;
        }
        t = fminf(t, __uint_as_float(a));
    }
    uint32_t base = 0;
    uint32_t pos[4];
    bool keep[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        keep[u] = v[u] && lo[u] <= t;
        const uint64_t bm = __ballot(keep[u]);
        pos[u] = base + (uint32_t)__popcll(bm & lane_mask_lt());
        base += (uint32_t)__popcll(bm);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (keep[u] && pos[u] < kAffKeep) {
            cj[pos[u]] = j[u];
            clo[pos[u]] = lo[u];
            chi[pos[u]] = __uint_as_float(hb[u]);
        }
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
        *T = t;
        if (base > kAffKeep) *over = 1u;  // the window outgrew the buffer: reported, not used
        *cnt = min(base, kAffKeep);
    }
}

// Workgroup: kAffWaves waves, 16 query rows each, whose FP16 operands for the whole K stay in
// registers (NCH chunks of 128 halfs: 16 NCH VGPRs); the columns stream through LDS in tiles of
// 64 x 128 halfs, double-buffered: the next chunk's global loads are in flight while the
// current one feeds the MFMAs.
constexpr int kAffWaves = 8;
constexpr int kAffThreads = kAffWaves * kWave;
constexpr int kAffRowsWG = 16 * kAffWaves;
constexpr int kAffChunkLoads = kAffCols * (kAffKC / 8) / kAffThreads;  // 16-byte loads per thread
static_assert(kAffChunkLoads * kAffThreads == kAffCols * (kAffKC / 8), "chunk loads");

template <int NCH>
__global__ __launch_bounds__(kAffThreads) void k_aff_candidates(const _Float16* __restrict__ Xh,
                                                                const float* __restrict__ nh,
                                                                const float* __restrict__ ah,
                                                                const float* __restrict__ eb,
                                                                uint32_t n, uint32_t k,
                                                                uint32_t* __restrict__ cj,
                                                                float* __restrict__ clo,
                                                                float* __restrict__ chi,
                                                                uint32_t* __restrict__ ccount,
                                                                uint32_t* __restrict__ overflow) {
    constexpr uint32_t dp = NCH * kAffKC;
    constexpr uint32_t ldb = kAffKC + kAffPad;
    __shared__ __attribute__((aligned(16))) _Float16 Bs[2][kAffCols * ldb];
    __shared__ uint32_t s_cnt[kAffRowsWG], s_over[kAffRowsWG];
    __shared__ float s_T[kAffRowsWG], s_nh[kAffRowsWG], s_ah[kAffRowsWG], s_eb[kAffRowsWG];
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int r16 = lane & 15, q4 = lane >> 4;
    const uint32_t row0 = blockIdx.x * kAffRowsWG;
    // this lane's A operands: row 16 w + r16, k = 32 s + 8 q4 .. + 7 for every step s
    half8 A[NCH * (kAffKC / 32)];
    {
        const uint32_t ar = row0 + 16 * w + r16;
#pragma unroll
        for (int s = 0; s < NCH * (kAffKC / 32); ++s)
            A[s] = ar < n ? *(const half8*)(Xh + (size_t)ar * dp + 32 * s + 8 * q4) : half8{};
    }
    if (tid < kAffRowsWG) {
        const uint32_t i = row0 + tid;
        s_cnt[tid] = 0u;
        s_over[tid] = 0u;
        s_T[tid] = __uint_as_float(0x7F800000u);
        s_nh[tid] = i < n ? nh[i] : 0.f;
        s_ah[tid] = i < n ? ah[i] : 0.f;
        s_eb[tid] = i < n ? eb[i] : 0.f;
    }
    const float kdelta = 0x1p-24f * (float)(2 * dp + 4);
    const uint32_t ntiles = div_up_a(n, kAffCols);
    const uint32_t nchunks = ntiles * NCH;  // (tile, chunk) in order
    // chunk c = (tile c / NCH, K chunk c % NCH): 64 columns x 16 loads of 16 bytes
    half8 pre[kAffChunkLoads];
    auto fetch = [&](uint32_t c) {
        const uint32_t c0 = (c / NCH) * kAffCols, ch = c % NCH;
#pragma unroll
        for (int u = 0; u < kAffChunkLoads; ++u) {
            const int e = tid + u * kAffThreads;
            const int col = e >> 4, q = e & 15;
            pre[u] = c0 + col < n
                         ? *(const half8*)(Xh + (size_t)(c0 + col) * dp + ch * kAffKC + q * 8)
                         : half8{};
        }
    };
    auto stage = [&](int buf) {
#pragma unroll
        for (int u = 0; u < kAffChunkLoads; ++u) {
            const int e = tid + u * kAffThreads;
            *(half8*)(&Bs[buf][(e >> 4) * ldb + (e & 15) * 8]) = pre[u];
        }
    };
    fetch(0);
    stage(0);
    __syncthreads();
    f32x4 acc[4];
    for (uint32_t t = 0; t < ntiles; ++t) {
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) acc[cb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            const uint32_t c = t * NCH + ch;
            const int buf = c & 1;
            if (c + 1 < nchunks) fetch(c + 1);  // in flight during the MFMAs below
#pragma unroll
            for (int s = 0; s < kAffKC / 32; ++s) {
                const half8 a = A[ch * (kAffKC / 32) + s];
#pragma unroll
                for (int cb = 0; cb < 4; ++cb) {
                    const half8 b =
                        *(const half8*)(&Bs[buf][(16 * cb + r16) * ldb + s * 32 + 8 * q4]);
                    acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc[cb], 0, 0, 0);
                }
            }
            // the other buffer was last read in the previous chunk, before its barrier
            if (c + 1 < nchunks) stage(buf ^ 1);
            __syncthreads();
        }
        // tile done: acc[cb][e] = x^_i . x^_j, i = row0 + 16 w + 4 q4 + e, j = c0 + 16 cb + r16
        const uint32_t c0 = t * kAffCols;
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            const uint32_t j = c0 + 16 * cb + r16;
            const bool jv = j < n;
            const float nj = jv ? nh[j] : 0.f, aj = jv ? ah[j] : 0.f, ej = jv ? eb[j] : 0.f;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int il = 16 * w + 4 * q4 + e;
                const uint32_t i = row0 + il;
                const float sn = s_nh[il] + nj;
                const float d = sn - 2.0f * acc[cb][e];
                const float delta = (kdelta * s_ah[il] * aj + 0x1p-22f * sn) * (1.0f + 0x1p-20f);
                const float e2 = (s_eb[il] + ej) * (1.0f + 0x1p-20f);
                float lo = sqrtf(fmaxf(d - delta, 0.f)) * (1.0f - 0x1p-20f) - e2;
                lo -= fabsf(lo) * 0x1p-22f;
                const float hi =
                    (sqrtf(fmaxf(d + delta, 0.f)) * (1.0f + 0x1p-20f) + e2) * (1.0f + 0x1p-22f);
                if (jv && i < n && j != i && lo <= s_T[il]) {
                    const uint32_t p = atomicAdd(&s_cnt[il], 1u);  // < kAffCap: see kAffKeep
                    const size_t o = (size_t)i * kAffCap + p;
                    cj[o] = j;
                    clo[o] = lo;
                    chi[o] = hi;
                }
            }
        }
        __syncthreads();
        // rows whose buffer could overflow with the next tile are compacted, a wave per row
        for (int r = w; r < kAffRowsWG; r += kAffWaves) {
            const uint32_t i = row0 + r;
            if (i < n && s_cnt[r] > kAffKeep)
                aff_compact(cj + (size_t)i * kAffCap, clo + (size_t)i * kAffCap,
                            chi + (size_t)i * kAffCap, &s_cnt[r], &s_T[r], &s_over[r], k);
        }
        __syncthreads();
    }
    for (int r = w; r < kAffRowsWG; r += kAffWaves) {
        const uint32_t i = row0 + r;
        if (i >= n) continue;
        aff_compact(cj + (size_t)i * kAffCap, clo + (size_t)i * kAffCap, chi + (size_t)i * kAffCap,
                    &s_cnt[r], &s_T[r], &s_over[r], k);
        if (lane == 0) {
            ccount[i] = s_cnt[r];
            if (s_over[r] || s_cnt[r] < k) atomicAdd(overflow, 1u);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// exact selection: FP64 distances of the candidates from the FP32 inputs, the k smallest by
// (distance, index), ascending (one wave per row)
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_aff_refine(const float* __restrict__ X, uint32_t n,
                                                       uint32_t dim, uint32_t k,
                                                       const uint32_t* __restrict__ cj,
                                                       const uint32_t* __restrict__ ccount,
                                                       uint32_t* __restrict__ idx,
                                                       double* __restrict__ d2) {
    __shared__ double s_d[kWavesPerBlock][kAffKeep];
    __shared__ uint32_t s_j[kWavesPerBlock][kAffKeep];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t i = blockIdx.x * kWavesPerBlock + w;
    if (i >= n) return;
    const uint32_t c = min(ccount[i], kAffKeep);
    float xi[kAffMaxDimPad / kWave];
#pragma unroll
    for (int t = 0; t < kAffMaxDimPad / kWave; ++t) {
        const uint32_t e = lane + kWave * t;
        xi[t] = e < dim ? X[(size_t)i * dim + e] : 0.f;
    }
    for (uint32_t q = 0; q < c; ++q) {
        const uint32_t j = cj[(size_t)i * kAffCap + q];
        double s = 0.0;
#pragma unroll
        for (int t = 0; t < kAffMaxDimPad / kWave; ++t) {
            const uint32_t e = lane + kWave * t;
            if (e < dim) {
                const double df = (double)xi[t] - (double)X[(size_t)j * dim + e];
                s = fma(df, df, s);
            }
        }
        s = wave_sum_d(s);
        if (lane == 0) {
            s_d[w][q] = s;
            s_j[w][q] = j;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t q = lane; q < c; q += kWave) {
        const double dq = s_d[w][q];
        const uint32_t jq = s_j[w][q];
        uint32_t rank = 0;
        for (uint32_t p = 0; p < c; ++p) {
            const double dp_ = s_d[w][p];
            rank += (dp_ < dq || (dp_ == dq && s_j[w][p] < jq)) ? 1u : 0u;
        }
        if (rank < k) {
            idx[(size_t)i * k + rank] = jq;
            d2[(size_t)i * k + rank] = dq;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// bandwidths: p_{j|i} = exp(-beta_i (d2_ij - d2_i0)) / sum, beta_i bisected until the entropy
// (natural log) is within tol of log(perplexity) (one wave per row, FP64; the scheme of
// openTSNE's binary search: doubling / halving until bracketed, then midpoints)
// ---------------------------------------------------------------------------------------------
constexpr int kAffMaxK = 128;
__global__ __launch_bounds__(kBlock) void k_aff_bandwidth(const double* __restrict__ d2,
                                                          uint32_t n, uint32_t k, double target,
                                                          double tol, int max_iter,
                                                          double* __restrict__ P,
                                                          double* __restrict__ beta_out) {
    const int lane = threadIdx.x & 63;
    const uint32_t i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (i >= n) return;
    const double* row = d2 + (size_t)i * k;
    const double d0 = row[0];
    double dr[kAffMaxK / kWave];
    bool v[kAffMaxK / kWave];
#pragma unroll
    for (int u = 0; u < kAffMaxK / kWave; ++u) {
        const uint32_t q = lane + kWave * u;
        v[u] = q < k;
        dr[u] = v[u] ? row[q] - d0 : 0.0;
    }
    double beta = 1.0, lo = -__builtin_huge_val(), hi = __builtin_huge_val();
    double p[kAffMaxK / kWave];
    for (int it = 0; it < max_iter; ++it) {
        double e[kAffMaxK / kWave], s = 0.0;
#pragma unroll
        for (int u = 0; u < kAffMaxK / kWave; ++u) {
            e[u] = v[u] ? exp(-beta * dr[u]) : 0.0;
            s += e[u];
        }
        s = wave_sum_d(s);
        double h = 0.0;
#pragma unroll
        for (int u = 0; u < kAffMaxK / kWave; ++u) {
            p[u] = e[u] / s;
            h += dr[u] * p[u];
        }
        h = log(s) + beta * wave_sum_d(h);
        const double diff = h - target;
        if (fabs(diff) < tol) break;
        if (diff > 0.0) {  // entropy too high: sharpen
            lo = beta;
            beta = __builtin_isinf(hi) ? beta * 2.0 : 0.5 * (beta + hi);
        } else {
            hi = beta;
            beta = __builtin_isinf(lo) ? beta * 0.5 : 0.5 * (beta + lo);
        }
    }
#pragma unroll
    for (int u = 0; u < kAffMaxK / kWave; ++u)
        if (v[u]) P[(size_t)i * k + lane + kWave * u] = p[u];
    if (lane == 0) beta_out[i] = beta;
}

// ---------------------------------------------------------------------------------------------
// symmetrisation: entries (i, j, p_{j|i}) and (j, i, p_{j|i}) keyed by row * n + column, sorted
// by the project's radix sort (stable), equal keys (at most two) summed in FP64 in sorted order
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_sym_entries(const uint32_t* __restrict__ idx,
                                                        uint32_t n, uint32_t k,
                                                        uint64_t* __restrict__ keys,
                                                        uint32_t* __restrict__ vals) {
    const uint64_t nk = (uint64_t)n * k;
    const uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= nk) return;
    const uint64_t i = e / k, j = idx[e];
    keys[e] = i * n + j;
    keys[nk + e] = j * n + i;
    vals[e] = (uint32_t)e;
    vals[nk + e] = (uint32_t)(nk + e);
}

__global__ __launch_bounds__(kBlock) void k_sym_heads(const uint64_t* __restrict__ keys,
                                                      uint32_t m, uint32_t* __restrict__ head) {
    const uint32_t u = blockIdx.x * kBlock + threadIdx.x;
    if (u >= m) return;
    head[u] = (u == 0 || keys[u] != keys[u - 1]) ? 1u : 0u;
}

__global__ __launch_bounds__(kBlock) void k_sym_reduce(const uint64_t* __restrict__ keys,
                                                       const uint32_t* __restrict__ vals,
                                                       const uint32_t* __restrict__ pos,
                                                       uint32_t m, uint32_t n, uint32_t nk,
                                                       const double* __restrict__ P,
                                                       uint32_t* __restrict__ col,
                                                       float* __restrict__ val,
                                                       uint32_t* __restrict__ row_ptr,
                                                       const uint32_t* __restrict__ nnz) {
    const uint32_t u = blockIdx.x * kBlock + threadIdx.x;
    if (u >= m) return;
    const uint64_t key = keys[u];
    if (u > 0 && keys[u - 1] == key) return;  // not the first of its key
    const double two_n = 2.0 * (double)n;
    double s = P[vals[u] % nk] / two_n;
    if (u + 1 < m && keys[u + 1] == key) s += P[vals[u + 1] % nk] / two_n;
    const uint32_t q = pos[u];
    const uint32_t r = (uint32_t)(key / n);
    col[q] = (uint32_t)(key % n);
    val[q] = (float)s;
    if (u == 0 || (uint32_t)(keys[u - 1] / n) != r) row_ptr[r] = q;  // every row has entries
    if (u == 0) row_ptr[n] = *nnz;
}

// ---------------------------------------------------------------------------------------------
uint32_t aff_dim_padded(uint32_t dim) { return div_up_a(dim, kAffKC) * kAffKC; }
uint32_t aff_cap() { return kAffCap; }

void launch_aff_pack(hipStream_t s, const float* X, uint32_t n, uint32_t dim, float scale,
                     _Float16* Xh, float* nh, float* ah, float* eb) {
    k_aff_pack<<<div_up_a(n, kWavesPerBlock), kBlock, 0, s>>>(X, n, dim, aff_dim_padded(dim),
                                                              scale, Xh, nh, ah, eb);
}
void launch_aff_candidates(hipStream_t s, const _Float16* Xh, const float* nh, const float* ah,
                           const float* eb, uint32_t n, uint32_t dim, uint32_t k, uint32_t* cj,
                           float* clo, float* chi, uint32_t* ccount, uint32_t* overflow) {
    const uint32_t g = div_up_a(n, kAffRowsWG);
    switch (aff_dim_padded(dim) / kAffKC) {
#define BH_AFFC(C)                                                                                \
    case C:                                                                                       \
        k_aff_candidates<C><<<g, kAffThreads, 0, s>>>(Xh, nh, ah, eb, n, k, cj, clo, chi, ccount,  \
                                                      overflow);                                  \
        break
        BH_AFFC(1); BH_AFFC(2); BH_AFFC(3); BH_AFFC(4);
        BH_AFFC(5); BH_AFFC(6); BH_AFFC(7); BH_AFFC(8);
#undef BH_AFFC
        default: break;  // dim > 1024 is rejected by the caller
    }
}
void launch_aff_refine(hipStream_t s, const float* X, uint32_t n, uint32_t dim, uint32_t k,
                       const uint32_t* cj, const uint32_t* ccount, uint32_t* idx, double* d2) {
    k_aff_refine<<<div_up_a(n, kWavesPerBlock), kBlock, 0, s>>>(X, n, dim, k, cj, ccount, idx,
                                                                d2);
}
void launch_aff_bandwidth(hipStream_t s, const double* d2, uint32_t n, uint32_t k, double target,
                          double tol, int max_iter, double* P, double* beta) {
    k_aff_bandwidth<<<div_up_a(n, kWavesPerBlock), kBlock, 0, s>>>(d2, n, k, target, tol,
                                                                   max_iter, P, beta);
}
void launch_sym_entries(hipStream_t s, const uint32_t* idx, uint32_t n, uint32_t k,
                        uint64_t* keys, uint32_t* vals) {
    k_sym_entries<<<div_up_a((uint64_t)n * k, kBlock), kBlock, 0, s>>>(idx, n, k, keys, vals);
}
void launch_sym_heads(hipStream_t s, const uint64_t* keys, uint32_t m, uint32_t* head) {
    k_sym_heads<<<div_up_a(m, kBlock), kBlock, 0, s>>>(keys, m, head);
}
void launch_sym_reduce(hipStream_t s, const uint64_t* keys, const uint32_t* vals,
                       const uint32_t* pos, uint32_t m, uint32_t n, uint32_t nk, const double* P,
                       uint32_t* col, float* val, uint32_t* row_ptr, const uint32_t* nnz) {
    k_sym_reduce<<<div_up_a(m, kBlock), kBlock, 0, s>>>(keys, vals, pos, m, n, nk, P, col, val,
                                                        row_ptr, nnz);
}

}  // namespace bh
