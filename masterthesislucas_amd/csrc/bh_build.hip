// Tree build of the Barnes-Hut force step on gfx950:
//   bounds -> Morton codes -> LDS-staged LSD radix sort -> gather into sorted SoA ->
//   leaf levels -> node count scan -> pre-order node emission -> upward mass/COM/M2 pass.
//
// The tree is laid out in pre-order over Morton-sorted particles: every node covers a
// contiguous particle range, its first child (if any) is the next node, and `skip` points past
// its subtree. Nodes starting at sorted index i are the cells of levels
// lcp(i-1, i)+1 .. leaf_level(i) whose first particle is i, so the layout follows from one scan.
#include "bh_common.h"
#include "bh_launch.h"

namespace bh {

// ------------------------------------------------------------------------------------------
// bounds: FP32 min/max, the last block derives the root square and the quantiser
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_bounds(const float* __restrict__ xs,
                                                   const float* __restrict__ ys, uint32_t n,
                                                   float4* __restrict__ partials,
                                                   uint32_t* __restrict__ counter,
                                                   BuildParams* __restrict__ prm, uint32_t tree_mode,
                                                   int use_caller, double cminx, double cminy,
                                                   double cmaxx, double cmaxy) {
    __shared__ float4 red[kWavesPerBlock];
    __shared__ int last;
    float mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const float x = xs[i], y = ys[i];
        mnx = fminf(mnx, x);
        mny = fminf(mny, y);
        mxx = fmaxf(mxx, x);
        mxy = fmaxf(mxy, y);
    }
    for (int o = 32; o > 0; o >>= 1) {
        mnx = fminf(mnx, __shfl_xor(mnx, o, kWave));
        mny = fminf(mny, __shfl_xor(mny, o, kWave));
        mxx = fmaxf(mxx, __shfl_xor(mxx, o, kWave));
        mxy = fmaxf(mxy, __shfl_xor(mxy, o, kWave));
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = make_float4(mnx, mny, mxx, mxy);
    __syncthreads();
    if (threadIdx.x == 0) {
        float4 r = red[0];
        for (int k = 1; k < kWavesPerBlock; ++k) {
            r.x = fminf(r.x, red[k].x);
            r.y = fminf(r.y, red[k].y);
            r.z = fmaxf(r.z, red[k].z);
            r.w = fmaxf(r.w, red[k].w);
        }
        partials[blockIdx.x] = r;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t t = atomicAdd(counter, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last || threadIdx.x != 0) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    float4 r = make_float4(INFINITY, INFINITY, -INFINITY, -INFINITY);
    for (uint32_t b = 0; b < gridDim.x; ++b) {
        float4 p;
        p.x = __hip_atomic_load(&partials[b].x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        p.y = __hip_atomic_load(&partials[b].y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        p.z = __hip_atomic_load(&partials[b].z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        p.w = __hip_atomic_load(&partials[b].w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        r.x = fminf(r.x, p.x);
        r.y = fminf(r.y, p.y);
        r.z = fmaxf(r.z, p.z);
        r.w = fmaxf(r.w, p.w);
    }
    *counter = 0u;  // ready for the next build (no memset node)
    uint64_t mnx64, mny64, mxx64, mxy64;
    if (use_caller) {
        mnx64 = bh_sfp_f64_bits(cminx);
        mny64 = bh_sfp_f64_bits(cminy);
        mxx64 = bh_sfp_f64_bits(cmaxx);
        mxy64 = bh_sfp_f64_bits(cmaxy);
    } else {
        mnx64 = bh_sfp_from_f32(r.x);
        mny64 = bh_sfp_from_f32(r.y);
        mxx64 = bh_sfp_from_f32(r.z);
        mxy64 = bh_sfp_from_f32(r.w);
    }
    uint64_t side = (tree_mode == kTreeQuadTree) ? bh_sfp_quadtree_side(mnx64, mny64, mxx64, mxy64)
                                                 : bh_sfp_morton_extent(mnx64, mny64, mxx64, mxy64);
    if ((side & ~BH_SFP_SIGN) == 0) side = 0x3FF0000000000000ull;  // all points coincide: side 1
    prm->minx = r.x;
    prm->miny = r.y;
    prm->maxx = r.z;
    prm->maxy = r.w;
    prm->minx64 = mnx64;
    prm->miny64 = mny64;
    prm->neg_minx64 = bh_sfp_neg(mnx64);
    prm->neg_miny64 = bh_sfp_neg(mny64);
    prm->side64 = side;
    prm->quant = bh_sfp_quant_prepare(side);
    prm->side = bh_sfp_to_f32(side);
    prm->n = n;
    prm->max_level = 0;
    prm->overfull_leaves = 0;
}

// ------------------------------------------------------------------------------------------
// Morton codes: bit-exact with the reference's FP64 rule via integer emulation
// ------------------------------------------------------------------------------------------
template <typename K>
__global__ __launch_bounds__(kBlock) void k_morton(const float* __restrict__ xs,
                                                   const float* __restrict__ ys, uint32_t n,
                                                   const BuildParams* __restrict__ prm, int B,
                                                   K* __restrict__ keys,
                                                   uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const bh_sfp_quant Q = prm->quant;
    const uint64_t dx = bh_sfp_add(bh_sfp_from_f32(xs[i]), prm->neg_minx64);
    const uint64_t dy = bh_sfp_add(bh_sfp_from_f32(ys[i]), prm->neg_miny64);
    const uint64_t qx = bh_sfp_quantize(dx, Q, B);
    const uint64_t qy = bh_sfp_quantize(dy, Q, B);
    K code;
    if (sizeof(K) == 4)
        code = (K)((bh_spread16((uint32_t)qy) << 1) | bh_spread16((uint32_t)qx));
    else
        code = (K)((bh_spread32(qy) << 1) | bh_spread32(qx));
    keys[i] = code;
    vals[i] = i;
}

// ------------------------------------------------------------------------------------------
// device-wide exclusive scan of uint32 (reduce / scan partials / downsweep)
// ------------------------------------------------------------------------------------------
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[kWavesPerBlock];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(x, o, kWave);
        if (lane >= o) x += t;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int k = 0; k < kWavesPerBlock; ++k) {
        const uint32_t s = wsum[k];
        if (k < w) off += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* __restrict__ in, uint32_t n,
                                                        uint32_t* __restrict__ partials) {
    const uint32_t base = blockIdx.x * kScanTile;
    uint32_t s = 0;
    for (int r = 0; r < kScanItems; ++r) {
        const uint32_t i = base + r * kBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t tot;
    block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBlock) void k_scan_partials(uint32_t* __restrict__ partials,
                                                          uint32_t nb, uint32_t* __restrict__ total) {
    const uint32_t per = (nb + kBlock - 1) / kBlock;
    const uint32_t lo = threadIdx.x * per;
    const uint32_t hi = min(lo + per, nb);
    uint32_t s = 0;
    for (uint32_t i = lo; i < hi; ++i) s += partials[i];
    uint32_t tot;
    uint32_t run = block_exclusive_scan(s, &tot);
    for (uint32_t i = lo; i < hi; ++i) {
        const uint32_t v = partials[i];
        partials[i] = run;
        run += v;
    }
    if (threadIdx.x == 0 && total) *total = tot;
}

__global__ __launch_bounds__(kBlock) void k_scan_down(const uint32_t* __restrict__ in,
                                                      uint32_t* __restrict__ out, uint32_t n,
                                                      const uint32_t* __restrict__ partials) {
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        v[r] = (base + r < n) ? in[base + r] : 0u;
        s += v[r];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(s, &tot) + partials[blockIdx.x];
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        if (base + r < n) out[base + r] = run;
        run += v[r];
    }
}

// ------------------------------------------------------------------------------------------
// LSD radix sort, 8-bit digits. Per pass: per-tile digit histogram -> device scan (digit-major)
// -> stable scatter. The scatter ranks keys inside each wave with ballot match masks, then
// stages the tile in LDS in digit order so global writes come out as contiguous digit runs.
// ------------------------------------------------------------------------------------------
constexpr int kSortItems = 16;                      // keys per lane
constexpr int kSortTile = kBlock * kSortItems;      // 4096 keys per block
constexpr int kSortWaveKeys = kWave * kSortItems;   // 1024 keys per wave

template <typename K>
__global__ __launch_bounds__(kBlock) void k_radix_hist(const K* __restrict__ keys, uint32_t n,
                                                       int shift, uint32_t nblocks,
                                                       uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kWavesPerBlock][256];
    const int w = threadIdx.x >> 6;
    for (int d = threadIdx.x; d < kWavesPerBlock * 256; d += kBlock) (&h[0][0])[d] = 0u;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t i = base + r * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[w][(uint32_t)(keys[i] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    const int d = threadIdx.x;
    uint32_t s = 0;
    for (int k = 0; k < kWavesPerBlock; ++k) s += h[k][d];
    hist[(uint32_t)d * nblocks + blockIdx.x] = s;
}

template <typename K>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const K* __restrict__ keys_in,
                                                          const uint32_t* __restrict__ vals_in,
                                                          K* __restrict__ keys_out,
                                                          uint32_t* __restrict__ vals_out,
                                                          uint32_t n, int shift, uint32_t nblocks,
                                                          const uint32_t* __restrict__ offsets) {
    __shared__ K s_keys[kSortTile];
    __shared__ uint32_t s_vals[kSortTile];
    __shared__ uint32_t s_wave[kWavesPerBlock][256];  // wave-local counts, then wave offsets
    __shared__ uint32_t s_start[256];                  // tile-local start of each digit
    __shared__ uint32_t s_base[256];                   // global start of each digit for this tile
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int d = threadIdx.x; d < kWavesPerBlock * 256; d += kBlock) (&s_wave[0][0])[d] = 0u;
    __syncthreads();

    const uint32_t wbase = blockIdx.x * kSortTile + w * kSortWaveKeys;
    const uint64_t lt = lane_mask_lt();
    K key[kSortItems];
    uint32_t val[kSortItems];
    uint32_t rank[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t i = wbase + r * kWave + lane;
        const bool valid = i < n;
        key[r] = valid ? keys_in[i] : (K)0;
        val[r] = valid ? vals_in[i] : 0u;
        const uint32_t d = (uint32_t)(key[r] >> shift) & 0xFFu;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t bits = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bits : ~bits;
        }
        const uint32_t before = s_wave[w][d];
        rank[r] = before + (uint32_t)__popcll(peers & lt);
        const bool leader = valid && ((peers & lt) == 0);
        if (leader) s_wave[w][d] = before + (uint32_t)__popcll(peers);
        if (!valid) rank[r] = 0xFFFFFFFFu;
    }
    __syncthreads();
    {
        const int d = threadIdx.x;
        uint32_t run = 0;
        for (int k = 0; k < kWavesPerBlock; ++k) {
            const uint32_t c = s_wave[k][d];
            s_wave[k][d] = run;
            run += c;
        }
        uint32_t tot;
        const uint32_t start = block_exclusive_scan(run, &tot);
        s_start[d] = start;
        s_base[d] = offsets[(uint32_t)d * nblocks + blockIdx.x];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        if (rank[r] == 0xFFFFFFFFu) continue;
        const uint32_t d = (uint32_t)(key[r] >> shift) & 0xFFu;
        const uint32_t pos = s_start[d] + s_wave[w][d] + rank[r];
        s_keys[pos] = key[r];
        s_vals[pos] = val[r];
    }
    __syncthreads();
    const uint32_t tile_n = min((uint32_t)kSortTile, n - blockIdx.x * kSortTile);
    for (uint32_t p = threadIdx.x; p < tile_n; p += kBlock) {
        const K k = s_keys[p];
        const uint32_t d = (uint32_t)(k >> shift) & 0xFFu;
        const uint32_t g = s_base[d] + (p - s_start[d]);
        keys_out[g] = k;
        vals_out[g] = s_vals[p];
    }
}

// ------------------------------------------------------------------------------------------
// gather the particle state into sorted order
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_gather(const uint32_t* __restrict__ perm, uint32_t n,
                                                   const float* __restrict__ x,
                                                   const float* __restrict__ y,
                                                   const float* __restrict__ m,
                                                   const float* __restrict__ vx,
                                                   const float* __restrict__ vy,
                                                   const uint32_t* __restrict__ id,
                                                   float* __restrict__ xo, float* __restrict__ yo,
                                                   float* __restrict__ mo, float* __restrict__ vxo,
                                                   float* __restrict__ vyo,
                                                   uint32_t* __restrict__ ido,
                                                   float4* __restrict__ pos4) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = perm[i];
    const float px = x[j], py = y[j], pm = m[j];
    xo[i] = px;
    yo[i] = py;
    mo[i] = pm;
    if (vx) {
        vxo[i] = vx[j];
        vyo[i] = vy[j];
    }
    ido[i] = id[j];
    pos4[i] = make_float4(px, py, pm, 0.0f);
}

// ------------------------------------------------------------------------------------------
// tree: leaf level per particle, node count per particle
// ------------------------------------------------------------------------------------------
template <typename K>
__global__ __launch_bounds__(kBlock) void k_tree_levels(const K* __restrict__ codes, uint32_t n,
                                                        int B, int cap, int fixed_depth,
                                                        uint8_t* __restrict__ leaf_lvl,
                                                        uint8_t* __restrict__ lcp1,
                                                        uint32_t* __restrict__ cnt,
                                                        BuildParams* __restrict__ prm) {
    __shared__ K s_code[kBlock + 2 * kMaxLeafCap];
    __shared__ int8_t s_w[kBlock + kMaxLeafCap];
    const int64_t b0 = (int64_t)blockIdx.x * kBlock;
    const int adaptive = fixed_depth < 0;
    if (adaptive) {
        const int span = kBlock + 2 * cap;
        for (int t = threadIdx.x; t < span; t += kBlock) {
            const int64_t j = b0 - cap + t;
            s_code[t] = (j >= 0 && j < (int64_t)n) ? codes[j] : (K)0;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < kBlock + cap; t += kBlock) {
            const int64_t j = b0 - cap + t;
            int8_t wv = -1;
            if (j >= 0 && j + cap < (int64_t)n) wv = (int8_t)lcp_level(s_code[t], s_code[t + cap], B);
            s_w[t] = wv;
        }
        __syncthreads();
    }
    const int64_t i = b0 + threadIdx.x;
    if (i >= (int64_t)n) return;
    int leaf;
    bool overfull = false;
    if (adaptive) {
        int m = -1;
        for (int t = 0; t <= cap; ++t) m = max(m, (int)s_w[threadIdx.x + t]);
        leaf = m + 1;
        if (leaf > B) {  // more than `cap` particles share the finest cell
            leaf = B;
            overfull = true;
        }
    } else {
        leaf = fixed_depth;
    }
    const int lp = (i == 0) ? -1 : lcp_level(codes[i - 1], codes[i], B);
    const uint32_t c = (leaf > lp) ? (uint32_t)(leaf - lp) : 0u;
    leaf_lvl[i] = (uint8_t)leaf;
    lcp1[i] = (uint8_t)(lp + 1);
    cnt[i] = c;
    if (c) {
        atomicMax(&prm->max_level, (uint32_t)leaf);
        if (overfull) atomicAdd(&prm->overfull_leaves, 1u);
    }
}

// first j >= lo with codes[j] > hi_key (codes ascending), galloping then binary search
template <typename K>
__device__ __forceinline__ uint32_t upper_from(const K* __restrict__ codes, uint32_t n, K hi_key,
                                               uint32_t lo) {
    uint32_t step = 1, hi = lo;
    while (hi < n && codes[hi] <= hi_key) {
        lo = hi + 1;
        hi = lo + step;
        step <<= 1;
    }
    if (hi > n) hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (codes[mid] <= hi_key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// first j <= hi with codes[j] >= lo_key (codes ascending), galloping backwards
template <typename K>
__device__ __forceinline__ uint32_t lower_back(const K* __restrict__ codes, K lo_key, uint32_t hi) {
    uint32_t step = 1, lo = hi;
    while (lo > 0 && codes[lo - 1] >= lo_key) {
        hi = lo - 1;
        lo = (hi > step) ? hi - step : 0;
        step <<= 1;
    }
    // answer in [lo, hi]: first index with codes[j] >= lo_key
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (codes[mid] >= lo_key)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

template <typename K>
__global__ __launch_bounds__(kBlock) void k_tree_emit(const K* __restrict__ codes, uint32_t n, int B,
                                                      const uint8_t* __restrict__ leaf_lvl,
                                                      const uint8_t* __restrict__ lcp1,
                                                      const uint32_t* __restrict__ cnt,
                                                      const uint32_t* __restrict__ noff,
                                                      const BuildParams* __restrict__ prm,
                                                      int single_child_pp, Node* __restrict__ nodes,
                                                      uint32_t* __restrict__ parent,
                                                      uint32_t* __restrict__ arrived,
                                                      uint32_t cap_nodes) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = cnt[i];
    if (!c) return;
    const uint32_t total = prm->n_nodes;
    const float side = prm->side;
    const int lp = (int)lcp1[i] - 1;
    const int leaf = leaf_lvl[i];
    const uint32_t base = noff[i];
    const K ci = codes[i];
    uint32_t e = i + 1;
    for (int t = (int)c - 1; t >= 0; --t) {
        const int l = lp + 1 + t;
        const int sh = 2 * (B - l);
        if (l == 0) {
            e = n;
        } else {
            const K hk = ci | (K)(((K)1 << sh) - 1);
            e = upper_from(codes, n, hk, e);
        }
        const uint32_t k = base + (uint32_t)t;
        if (k >= cap_nodes) continue;  // cannot happen for the capacity the context allocates
        uint32_t meta = (uint32_t)l;
        if (l == leaf) {
            meta |= kMetaLeaf | kMetaPP;
        } else if (single_child_pp) {
            const int shc = 2 * (B - l - 1);
            const uint32_t d0 = (uint32_t)(ci >> shc) & 3u;
            const uint32_t d1 = (uint32_t)(codes[e - 1] >> shc) & 3u;
            if (d0 == d1) meta |= kMetaPP;
        }
        const float sl = ldexpf(side, -l);
        Node* nd = nodes + k;
        nd->l2 = sl * sl;
        nd->skip = (e < n) ? noff[e] : total;
        nd->first = i;
        nd->count = e - i;
        nd->meta = meta;
        arrived[k] = 0u;
        uint32_t par;
        if (l == 0) {
            par = 0xFFFFFFFFu;
        } else if (t > 0) {
            par = k - 1;
        } else {
            const int pl = l - 1;
            uint32_t ip = 0;
            if (pl > 0) {
                const int psh = 2 * (B - pl);
                const K lk = ci & ~(K)(((K)1 << psh) - 1);
                ip = lower_back(codes, lk, i);
            }
            par = noff[ip] + (uint32_t)(pl - ((int)lcp1[ip]));
        }
        parent[k] = par;
    }
}

// ------------------------------------------------------------------------------------------
// upward pass: leaves sum their particles; the last child to finish combines its parent
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_tree_upward(const float4* __restrict__ pos4,
                                                        Node* __restrict__ nodes,
                                                        const uint32_t* __restrict__ parent,
                                                        uint32_t* __restrict__ arrived,
                                                        float4* __restrict__ m2,
                                                        const BuildParams* __restrict__ prm,
                                                        uint32_t cap_nodes) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= cap_nodes || k >= prm->n_nodes) return;
    Node* nd = nodes + k;
    if (!(nd->meta & kMetaLeaf)) return;
    const uint32_t f = nd->first, c = nd->count;
    const float4 p0 = pos4[f];
    float M = 0.f, sx = 0.f, sy = 0.f;
    for (uint32_t j = f; j < f + c; ++j) {
        const float4 p = pos4[j];
        M += p.z;
        sx = fmaf(p.z, p.x - p0.x, sx);
        sy = fmaf(p.z, p.y - p0.y, sy);
    }
    const float inv = (M != 0.f) ? 1.0f / M : 0.f;
    float cx = p0.x + sx * inv, cy = p0.y + sy * inv;
    nd->cx = cx;
    nd->cy = cy;
    nd->mass = M;
    if (m2) {
        float a = 0.f, b = 0.f, d = 0.f;
        for (uint32_t j = f; j < f + c; ++j) {
            const float4 p = pos4[j];
            const float rx = p.x - cx, ry = p.y - cy;
            a = fmaf(p.z * rx, rx, a);
            b = fmaf(p.z * rx, ry, b);
            d = fmaf(p.z * ry, ry, d);
        }
        m2[k] = make_float4(a, b, d, 0.f);
    }
    uint32_t cur = k, cur_count = c;
    while (true) {
        const uint32_t p = parent[cur];
        if (p == 0xFFFFFFFFu) break;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t old = atomicAdd(&arrived[p], cur_count);
        const uint32_t pc = nodes[p].count;
        if (old + cur_count != pc) break;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const uint32_t end = nodes[p].skip;
        // children: p+1, then follow skip pointers until the parent's subtree ends
        uint32_t ch = p + 1;
        const float ox = __hip_atomic_load(&nodes[ch].cx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const float oy = __hip_atomic_load(&nodes[ch].cy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        float PM = 0.f, px = 0.f, py = 0.f;
        while (ch < end) {
            const float cm = __hip_atomic_load(&nodes[ch].mass, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const float ccx = __hip_atomic_load(&nodes[ch].cx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const float ccy = __hip_atomic_load(&nodes[ch].cy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            PM += cm;
            px = fmaf(cm, ccx - ox, px);
            py = fmaf(cm, ccy - oy, py);
            ch = nodes[ch].skip;
        }
        const float pinv = (PM != 0.f) ? 1.0f / PM : 0.f;
        const float pcx = ox + px * pinv, pcy = oy + py * pinv;
        nodes[p].cx = pcx;
        nodes[p].cy = pcy;
        nodes[p].mass = PM;
        if (m2) {
            float a = 0.f, b = 0.f, d = 0.f;
            ch = p + 1;
            while (ch < end) {
                const float cm = __hip_atomic_load(&nodes[ch].mass, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const float rx = __hip_atomic_load(&nodes[ch].cx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - pcx;
                const float ry = __hip_atomic_load(&nodes[ch].cy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - pcy;
                const float ca = __hip_atomic_load(&m2[ch].x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const float cb = __hip_atomic_load(&m2[ch].y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const float cd = __hip_atomic_load(&m2[ch].z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                a += ca + cm * rx * rx;
                b += cb + cm * rx * ry;
                d += cd + cm * ry * ry;
                ch = nodes[ch].skip;
            }
            m2[p] = make_float4(a, b, d, 0.f);
        }
        cur = p;
        cur_count = pc;
    }
}

// ------------------------------------------------------------------------------------------
// reference slot layout of the fixed-depth Morton tree (nBodySolverFMM_SYM_MORTON.h:117-151,
// 328-470): slot = levelIndex[l] + cell index, firstChildIndex = levelIndex[l+1] + 4 cell
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_iota(uint32_t* __restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = i;
}

__global__ __launch_bounds__(kBlock) void k_tree_slots(const uint32_t* __restrict__ codes,
                                                       const Node* __restrict__ nodes,
                                                       const BuildParams* __restrict__ prm,
                                                       uint32_t cap_nodes, int depth,
                                                       uint32_t* __restrict__ first_child,
                                                       uint32_t* __restrict__ first_particle,
                                                       uint32_t* __restrict__ count) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= cap_nodes || k >= prm->n_nodes) return;
    const Node nd = nodes[k];
    const int l = (int)(nd.meta & kMetaLevelMask);
    const uint32_t cell = (l == 0) ? 0u : (codes[nd.first] >> (2 * (16 - l)));
    const uint32_t li = (uint32_t)((((uint64_t)1 << (2 * l)) - 1) / 3);
    const uint32_t slot = li + cell;
    first_particle[slot] = nd.first;
    count[slot] = nd.count;
    first_child[slot] = (l < depth) ? (uint32_t)((((uint64_t)1 << (2 * (l + 1))) - 1) / 3) + 4u * cell : 0u;
}

// ------------------------------------------------------------------------------------------
// host launchers
// ------------------------------------------------------------------------------------------
static inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

uint32_t bounds_grid(uint32_t n) {
    uint32_t g = div_up(n, kBlock * 8);
    if (g < 1) g = 1;
    if (g > kMaxBoundsBlocks) g = kMaxBoundsBlocks;
    return g;
}

void launch_bounds(hipStream_t s, const float* xs, const float* ys, uint32_t n, float4* partials,
                   uint32_t* counter, BuildParams* prm, uint32_t tree_mode, const double* caller) {
    const uint32_t g = bounds_grid(n);
    k_bounds<<<g, kBlock, 0, s>>>(xs, ys, n, partials, counter, prm, tree_mode, caller ? 1 : 0,
                                  caller ? caller[0] : 0.0, caller ? caller[1] : 0.0,
                                  caller ? caller[2] : 0.0, caller ? caller[3] : 0.0);
}

void launch_morton(hipStream_t s, const float* xs, const float* ys, uint32_t n,
                   const BuildParams* prm, int B, void* keys, uint32_t* vals, bool wide) {
    const uint32_t g = div_up(n, kBlock);
    if (wide)
        k_morton<uint64_t><<<g, kBlock, 0, s>>>(xs, ys, n, prm, B, (uint64_t*)keys, vals);
    else
        k_morton<uint32_t><<<g, kBlock, 0, s>>>(xs, ys, n, prm, B, (uint32_t*)keys, vals);
}

uint32_t scan_partials_size(uint64_t n) { return div_up(n, kScanTile) + 1; }

void launch_exclusive_scan(hipStream_t s, const uint32_t* in, uint32_t* out, uint32_t n,
                           uint32_t* partials, uint32_t* total) {
    const uint32_t nb = div_up(n, kScanTile);
    k_scan_reduce<<<nb, kBlock, 0, s>>>(in, n, partials);
    k_scan_partials<<<1, kBlock, 0, s>>>(partials, nb, total);
    k_scan_down<<<nb, kBlock, 0, s>>>(in, out, n, partials);
}

uint32_t sort_blocks(uint32_t n) { return div_up(n, kSortTile); }

void launch_radix_sort(hipStream_t s, void* keys_a, uint32_t* vals_a, void* keys_b,
                       uint32_t* vals_b, uint32_t n, int key_bits, bool wide, uint32_t* hist,
                       uint32_t* scan_partials, bool* result_in_b) {
    const uint32_t nb = sort_blocks(n);
    const uint32_t nh = nb * 256u;
    const int passes = (key_bits + 7) / 8;
    void* kin = keys_a;
    uint32_t* vin = vals_a;
    void* kout = keys_b;
    uint32_t* vout = vals_b;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        if (wide)
            k_radix_hist<uint64_t><<<nb, kBlock, 0, s>>>((const uint64_t*)kin, n, shift, nb, hist);
        else
            k_radix_hist<uint32_t><<<nb, kBlock, 0, s>>>((const uint32_t*)kin, n, shift, nb, hist);
        launch_exclusive_scan(s, hist, hist, nh, scan_partials, nullptr);
        if (wide)
            k_radix_scatter<uint64_t><<<nb, kBlock, 0, s>>>((const uint64_t*)kin, vin,
                                                            (uint64_t*)kout, vout, n, shift, nb, hist);
        else
            k_radix_scatter<uint32_t><<<nb, kBlock, 0, s>>>((const uint32_t*)kin, vin,
                                                            (uint32_t*)kout, vout, n, shift, nb, hist);
        void* tk = kin; kin = kout; kout = tk;
        uint32_t* tv = vin; vin = vout; vout = tv;
    }
    *result_in_b = (passes & 1) != 0;
}

void launch_gather(hipStream_t s, const uint32_t* perm, uint32_t n, const float* x, const float* y,
                   const float* m, const float* vx, const float* vy, const uint32_t* id, float* xo,
                   float* yo, float* mo, float* vxo, float* vyo, uint32_t* ido, float4* pos4) {
    k_gather<<<div_up(n, kBlock), kBlock, 0, s>>>(perm, n, x, y, m, vx, vy, id, xo, yo, mo, vxo,
                                                  vyo, ido, pos4);
}

void launch_tree(hipStream_t s, const void* codes, bool wide, uint32_t n, int B, int cap,
                 int fixed_depth, int single_child_pp, uint8_t* leaf_lvl, uint8_t* lcp1,
                 uint32_t* cnt, uint32_t* noff, uint32_t* scan_partials, BuildParams* prm,
                 Node* nodes, uint32_t* parent, uint32_t* arrived, float4* m2, uint32_t cap_nodes) {
    const uint32_t g = div_up(n, kBlock);
    if (wide)
        k_tree_levels<uint64_t><<<g, kBlock, 0, s>>>((const uint64_t*)codes, n, B, cap,
                                                     fixed_depth, leaf_lvl, lcp1, cnt, prm);
    else
        k_tree_levels<uint32_t><<<g, kBlock, 0, s>>>((const uint32_t*)codes, n, B, cap,
                                                     fixed_depth, leaf_lvl, lcp1, cnt, prm);
    launch_exclusive_scan(s, cnt, noff, n, scan_partials, &prm->n_nodes);
    if (wide)
        k_tree_emit<uint64_t><<<g, kBlock, 0, s>>>((const uint64_t*)codes, n, B, leaf_lvl, lcp1,
                                                   cnt, noff, prm, single_child_pp, nodes, parent,
                                                   arrived, cap_nodes);
    else
        k_tree_emit<uint32_t><<<g, kBlock, 0, s>>>((const uint32_t*)codes, n, B, leaf_lvl, lcp1,
                                                   cnt, noff, prm, single_child_pp, nodes, parent,
                                                   arrived, cap_nodes);
}

void launch_upward(hipStream_t s, const float4* pos4, Node* nodes, const uint32_t* parent,
                   uint32_t* arrived, float4* m2, const BuildParams* prm, uint32_t cap_nodes) {
    k_tree_upward<<<div_up(cap_nodes, kBlock), kBlock, 0, s>>>(pos4, nodes, parent, arrived, m2,
                                                                prm, cap_nodes);
}

void launch_fill_u32(hipStream_t s, uint32_t* p, uint64_t n, uint32_t v) {
    if (n) k_fill_u32<<<div_up(n, kBlock), kBlock, 0, s>>>(p, n, v);
}

void launch_fill_f32(hipStream_t s, float* p, uint64_t n, float v) {
    union {
        float f;
        uint32_t u;
    } c;
    c.f = v;
    launch_fill_u32(s, (uint32_t*)p, n, c.u);
}

void launch_iota(hipStream_t s, uint32_t* p, uint32_t n) {
    if (n) k_iota<<<div_up(n, kBlock), kBlock, 0, s>>>(p, n);
}

void launch_tree_slots(hipStream_t s, const uint32_t* codes, const Node* nodes,
                       const BuildParams* prm, uint32_t cap_nodes, int depth, uint32_t* first_child,
                       uint32_t* first_particle, uint32_t* count, uint64_t slots) {
    k_fill_u32<<<div_up(slots, kBlock), kBlock, 0, s>>>(first_child, slots, 0u);
    k_fill_u32<<<div_up(slots, kBlock), kBlock, 0, s>>>(first_particle, slots, 0u);
    k_fill_u32<<<div_up(slots, kBlock), kBlock, 0, s>>>(count, slots, 0u);
    k_tree_slots<<<div_up(cap_nodes, kBlock), kBlock, 0, s>>>(codes, nodes, prm, cap_nodes, depth,
                                                              first_child, first_particle, count);
}

}  // namespace bh
