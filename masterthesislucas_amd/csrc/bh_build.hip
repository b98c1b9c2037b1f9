// Tree build of the Barnes-Hut force step on gfx950:
//   bounds -> Morton codes -> LDS-staged LSD radix sort -> gather into sorted SoA ->
//   leaf levels -> node count scan -> pre-order node emission -> upward mass/COM/M2 pass.
//
// The tree is laid out in pre-order over Morton-sorted particles: every node covers a
// contiguous particle range, its first child (if any) is the next node, and `skip` points past
// its subtree. Nodes starting at sorted index i are the cells of levels
// lcp(i-1, i)+1 .. leaf_level(i) whose first particle is i, so the layout follows from one scan.
//
// No kernel here uses device-scope fences or same-address device atomics: on a multi-XCD
// part both are expensive (a device-scope fence writes back or invalidates the XCD's L2).
// Results pass between kernels, and within a kernel only through LDS.
#include "bh_common.h"
#include "bh_launch.h"

namespace bh {

// ------------------------------------------------------------------------------------------
// bounds: FP32 min/max per block, then one block derives the root square and the quantiser
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float4 minmax4(float4 a, float4 b) {
    return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fmaxf(a.z, b.z), fmaxf(a.w, b.w));
}

__device__ __forceinline__ float4 block_minmax4(float4 r) {
    __shared__ float4 red[kWavesPerBlock];
    for (int o = 32; o > 0; o >>= 1) {
        r.x = fminf(r.x, __shfl_xor(r.x, o, kWave));
        r.y = fminf(r.y, __shfl_xor(r.y, o, kWave));
        r.z = fmaxf(r.z, __shfl_xor(r.z, o, kWave));
        r.w = fmaxf(r.w, __shfl_xor(r.w, o, kWave));
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = r;
    __syncthreads();
    r = red[0];
    for (int k = 1; k < kWavesPerBlock; ++k) r = minmax4(r, red[k]);
    __syncthreads();
    return r;
}

// One position as a bounds item. fminf/fmaxf drop NaN, so a NaN coordinate enters as -inf:
// the bounds are then finite exactly when every position is (an infinity stays infinite).
__device__ __forceinline__ float4 bounds_item(float x, float y) {
    const float xl = x == x ? x : -INFINITY, yl = y == y ? y : -INFINITY;
    return make_float4(xl, yl, x, y);
}

// status word 1: the build saw a non-finite position (read by the host after the build)
__device__ __forceinline__ void flag_nonfinite(float4 r, uint32_t* status) {
    if (!(isfinite(r.x) && isfinite(r.y) && isfinite(r.z) && isfinite(r.w))) status[1] = 1u;
}

__global__ __launch_bounds__(kBlock) void k_bounds_partial(const float* __restrict__ xs,
                                                           const float* __restrict__ ys,
                                                           uint32_t n,
                                                           float4* __restrict__ partials) {
    float4 r = make_float4(INFINITY, INFINITY, -INFINITY, -INFINITY);
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        r = minmax4(r, bounds_item(xs[i], ys[i]));
    r = block_minmax4(r);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

__device__ void fill_params(BuildParams* prm, float4 r, uint64_t mnx64, uint64_t mny64,
                            uint64_t mxx64, uint64_t mxy64, uint32_t tree_mode, uint32_t n);

__global__ __launch_bounds__(kBlock) void k_bounds_final(const float4* __restrict__ partials,
                                                         uint32_t nb, uint32_t n,
                                                         BuildParams* __restrict__ prm,
                                                         uint32_t tree_mode, int use_caller,
                                                         double cminx, double cminy,
                                                         double cmaxx, double cmaxy,
                                                         uint32_t* __restrict__ status) {
    float4 r = make_float4(INFINITY, INFINITY, -INFINITY, -INFINITY);
    for (uint32_t b = threadIdx.x; b < nb; b += kBlock) r = minmax4(r, partials[b]);
    r = block_minmax4(r);
    if (threadIdx.x != 0) return;
    flag_nonfinite(r, status);
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
    fill_params(prm, r, mnx64, mny64, mxx64, mxy64, tree_mode, n);
}

// batched systems: bounds per system from per-block partials over each system's blocks
__global__ __launch_bounds__(kBlock) void k_bounds_seg_partial(const float* __restrict__ xs,
                                                               const float* __restrict__ ys,
                                                               const uint2* __restrict__ blk,
                                                               float4* __restrict__ partials) {
    const uint2 bk = blk[blockIdx.x];  // (first particle, count)
    float4 r = make_float4(INFINITY, INFINITY, -INFINITY, -INFINITY);
    for (uint32_t i = threadIdx.x; i < bk.y; i += kBlock)
        r = minmax4(r, bounds_item(xs[bk.x + i], ys[bk.x + i]));
    r = block_minmax4(r);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// one block per system: partials [blk_off[s], blk_off[s+1]) -> sys_prm[s]
__global__ __launch_bounds__(kBlock) void k_bounds_seg_final(const float4* __restrict__ partials,
                                                             const uint32_t* __restrict__ blk_off,
                                                             const uint32_t* __restrict__ sys_off,
                                                             BuildParams* __restrict__ sys_prm,
                                                             BuildParams* __restrict__ prm,
                                                             uint32_t* __restrict__ status) {
    const uint32_t s = blockIdx.x;
    float4 r = make_float4(INFINITY, INFINITY, -INFINITY, -INFINITY);
    for (uint32_t b = blk_off[s] + threadIdx.x; b < blk_off[s + 1]; b += kBlock)
        r = minmax4(r, partials[b]);
    r = block_minmax4(r);
    if (threadIdx.x != 0) return;
    flag_nonfinite(r, status);
    fill_params(sys_prm + s, r, bh_sfp_from_f32(r.x), bh_sfp_from_f32(r.y), bh_sfp_from_f32(r.z),
                bh_sfp_from_f32(r.w), kTreeQuadTree, sys_off[s + 1] - sys_off[s]);
    if (s == 0) {
        prm->n = sys_off[gridDim.x];
        prm->max_level = 0;
        prm->overfull_leaves = 0;
    }
}

__device__ void fill_params(BuildParams* prm, float4 r, uint64_t mnx64, uint64_t mny64,
                            uint64_t mxx64, uint64_t mxy64, uint32_t tree_mode, uint32_t n) {
    uint64_t side;
    if (tree_mode == kTreeQuadTree)
        side = bh_sfp_quadtree_side(mnx64, mny64, mxx64, mxy64);
    else if (tree_mode == kTreeGpuRef)  // createRoot: the larger FP32 extent (.cu:78-94)
        side = bh_sfp_from_f32(fmaxf(r.z - r.x, r.w - r.y));
    else
        side = bh_sfp_morton_extent(mnx64, mny64, mxx64, mxy64);
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

// The tree of the reference's GPU step (NBodySolverGpuBH: createRoot, createChildren,
// nBodySolverGpuBH.cu:78-200): root centred on the bounding box (FP32 midpoint) with side the
// larger FP32 extent; every cell splits at its centre, a point equal to the centre coordinate
// going low; child centres c +- 0.25 l and side 0.5 l in FP32 as there. Each point descends
// `depth` levels by those comparisons; the code holds its path two bits per level, x on the
// even bit as in the Morton codes (the reference's child order A, B, C, D is (x, y) =
// (low, low), (low, high), (high, low), (high, high); only the storage order differs), in the
// top 2 depth bits of the 32-bit key.
__global__ __launch_bounds__(kBlock) void k_morton_gpuref(const float* __restrict__ xs,
                                                          const float* __restrict__ ys,
                                                          uint32_t n,
                                                          const BuildParams* __restrict__ prm,
                                                          int depth, uint32_t* __restrict__ keys,
                                                          uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float mnx = prm->minx, mny = prm->miny, mxx = prm->maxx, mxy = prm->maxy;
    float len = fmaxf(mxx - mnx, mxy - mny);
    float cx = (mnx + mxx) * 0.5f, cy = (mny + mxy) * 0.5f;
    const float x = xs[i], y = ys[i];
    uint32_t code = 0;
    for (int l = 0; l < depth; ++l) {
        const bool xh = !(x <= cx), yh = !(y <= cy);
        code = (code << 2) | ((uint32_t)yh << 1) | (uint32_t)xh;
        const float q = 0.25f * len;
        cx = cx + (xh ? q : -q);
        cy = cy + (yh ? q : -q);
        len = 0.5f * len;
    }
    keys[i] = depth > 0 ? code << (32 - 2 * depth) : 0u;
    vals[i] = i;
}

// batched systems: key = system << 2B | code, with each system's own root square
__global__ __launch_bounds__(kBlock) void k_morton_seg(const float* __restrict__ xs,
                                                       const float* __restrict__ ys, uint32_t n,
                                                       const uint32_t* __restrict__ sys_off,
                                                       uint32_t nsys,
                                                       const BuildParams* __restrict__ sys_prm,
                                                       int B, uint64_t* __restrict__ keys,
                                                       uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t lo = 0, hi = nsys;  // largest s with sys_off[s] <= i
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sys_off[mid] <= i) lo = mid; else hi = mid;
    }
    const BuildParams* prm = sys_prm + lo;
    const bh_sfp_quant Q = prm->quant;
    const uint64_t dx = bh_sfp_add(bh_sfp_from_f32(xs[i]), prm->neg_minx64);
    const uint64_t dy = bh_sfp_add(bh_sfp_from_f32(ys[i]), prm->neg_miny64);
    const uint64_t qx = bh_sfp_quantize(dx, Q, B);
    const uint64_t qy = bh_sfp_quantize(dy, Q, B);
    keys[i] = ((uint64_t)lo << (2 * B)) | (bh_spread32(qy) << 1) | bh_spread32(qx);
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

// the same for a block of THREADS threads
template <int THREADS>
__device__ __forceinline__ uint32_t block_exclusive_scan_t(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[THREADS / kWave];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int k = 0; k < THREADS / kWave; ++k) {
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
// LSD radix sort, 8-bit digits, one launch per pass after one counting launch ("one-sweep").
//   k_radix_counts: one block per 4096-key tile; the tile's digit counts for every pass (the
//     multiset of digits of a pass does not depend on the order) and the clearing of the tile's
//     look-back words of every pass.
//   k_radix_totals: per (pass, digit) the total over tiles.
//   k_radix_pass: a block takes the next tile ticket, ranks its keys inside each wave with ballot
//     match masks, publishes the tile's digit counts, and gets the tile's global digit offsets
//     by decoupled look-back over the tiles ticketed before it; it then stages the tile in LDS
//     in digit order so global writes come out as contiguous digit runs. The sort is stable.
// A look-back word holds a 2-bit state (0 not ready, 1 tile count, 2 inclusive prefix) above a
// 30-bit count, so one relaxed device-scope atomic carries both and no fence is needed. A block
// only waits on tiles whose tickets were taken before its own (their blocks are running), so
// every wait ends.
// ------------------------------------------------------------------------------------------
// Tiles of 4096 keys for large sorts, 2048 for sorts of at most kSortSmallMax keys (more
// blocks share a small sort: 70k keys 20% faster, while 1M keys are 14% slower with them).
// kSortItems* is the tile in keys per thread of a 256-thread block (the counting kernel); the
// pass kernel ranks a tile with 4 keys per lane (1024 threads, or 512 for small tiles), which
// measured faster than 8 or 16 keys per lane on fewer waves.
constexpr int kSortItemsLarge = 16, kSortItemsSmall = 8;
constexpr uint32_t kSortSmallMax = 1u << 18;  // sorts up to this size use the small tiles
__host__ __device__ constexpr int sort_tile(int items) { return kBlock * items; }
constexpr int kSortMaxPasses = 8;
constexpr uint32_t kLbCount = 1u << 30, kLbPrefix = 2u << 30, kLbMask = kLbCount - 1;
constexpr int kLbWindow = 8;  // look-back loads in flight per digit
static_assert(kBlock == 256, "one thread per digit");

template <typename K, int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS) void k_radix_counts(const K* __restrict__ keys, uint32_t n,
                                                          int passes, uint32_t ntiles,
                                                          uint32_t* __restrict__ counts,
                                                          uint32_t* __restrict__ status) {
    __shared__ uint32_t h[kSortMaxPasses][256];
    for (int i = threadIdx.x; i < kSortMaxPasses * 256; i += THREADS) (&h[0][0])[i] = 0u;
    for (int i = threadIdx.x; i < passes * 256; i += THREADS)
        status[((size_t)(i >> 8) * ntiles + blockIdx.x) * 256 + (i & 255)] = 0u;
    __syncthreads();
    const uint32_t base = blockIdx.x * (THREADS * ITEMS);
    K k[ITEMS];  // all loads issued before the first use
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const uint32_t i = base + r * THREADS + threadIdx.x;
        k[r] = i < n ? keys[i] : (K)0;
    }
    // A wave holds 64 consecutive keys; each run of equal digits among neighbouring lanes is
    // counted by its first lane with one add of the run length. Keys that arrive nearly sorted
    // (particles kept in the previous step's order) would otherwise send most of a wave's adds
    // to one LDS word, which serialises them.
    const int lane = threadIdx.x & 63;
    const uint64_t le = lane_mask_lt() | (1ull << lane);
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const bool valid = base + r * THREADS + threadIdx.x < n;  // invalid lanes are a suffix
        const K prev = (K)__shfl_up((unsigned long long)k[r], 1);
        const uint64_t stops = __ballot(!valid);
        for (int p = 0; p < passes; ++p) {
            const uint32_t d = (uint32_t)(k[r] >> (8 * p)) & 0xFFu;
            const bool head = valid && (lane == 0 || d != ((uint32_t)(prev >> (8 * p)) & 0xFFu));
            const uint64_t after = (__ballot(head) | stops) & ~le;
            const uint32_t end = after ? (uint32_t)__builtin_ctzll(after) : 64u;
            if (head) atomicAdd(&h[p][d], end - (uint32_t)lane);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * 256; i += THREADS)
        counts[(size_t)i * ntiles + blockIdx.x] = h[i >> 8][i & 255];
}

// one wave per (pass, digit) row of counts; block 0 also clears the pass tickets
__global__ __launch_bounds__(kBlock) void k_radix_totals(const uint32_t* __restrict__ counts,
                                                         uint32_t ntiles, int passes,
                                                         uint32_t* __restrict__ totals,
                                                         uint32_t* __restrict__ tickets) {
    if (blockIdx.x == 0 && threadIdx.x < kSortMaxPasses) tickets[threadIdx.x] = 0u;
    const int lane = threadIdx.x & 63;
    const uint32_t row = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (row >= (uint32_t)passes * 256u) return;
    unsigned long long s = 0;
    for (uint32_t j = lane; j < ntiles; j += kWave) s += counts[(size_t)row * ntiles + j];
    s = wave_sum_u64(s);
    if (lane == 0) totals[row] = (uint32_t)s;
}

template <typename K, int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS) void k_radix_pass(const K* __restrict__ keys_in,
                                                       const uint32_t* __restrict__ vals_in,
                                                       K* __restrict__ keys_out,
                                                       uint32_t* __restrict__ vals_out,
                                                       uint32_t n, int shift,
                                                       const uint32_t* __restrict__ totals,
                                                       uint32_t* __restrict__ status,
                                                       uint32_t* __restrict__ ticket,
                                                       uint32_t* __restrict__ err,
                                                       uint32_t spin_limit) {
    constexpr int kTile = THREADS * ITEMS, kWaveKeys = kWave * ITEMS, kW = THREADS / kWave;
    // After a give-up the keys of the next pass no longer match the digit totals, and placing
    // them could write past the end: every later block leaves at once (the build reports the
    // error and its gather falls back to the identity order).
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return;
    static_assert(THREADS >= 256, "one thread per digit");
    __shared__ K s_keys[kTile];
    __shared__ uint32_t s_vals[kTile];
    __shared__ uint32_t s_wave[kW][256];  // wave-local counts, then wave offsets
    __shared__ uint32_t s_start[256];                  // tile-local start of each digit
    __shared__ uint32_t s_base[256];                   // global start of each digit for this tile
    __shared__ uint32_t s_tile;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    for (int d = threadIdx.x; d < kW * 256; d += THREADS) (&s_wave[0][0])[d] = 0u;
    __syncthreads();
    const uint32_t tile = s_tile;

    const uint32_t wbase = tile * kTile + w * kWaveKeys;
    const uint64_t lt = lane_mask_lt();
    K key[ITEMS];
    uint32_t val[ITEMS];
    uint32_t rank[ITEMS];
    // all loads are issued before the ranking (whose LDS traffic would otherwise make each
    // item wait for its own load)
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const uint32_t i = wbase + r * kWave + lane;
        key[r] = i < n ? keys_in[i] : (K)0;
        val[r] = i < n ? vals_in[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const bool valid = wbase + r * kWave + lane < n;
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
        // threads 0..255 own one digit each; the scans take every thread of the block
        const int d = threadIdx.x & 255;
        const bool dig = threadIdx.x < 256;
        uint32_t cnt = 0;
        if (dig) {
            for (int k = 0; k < kW; ++k) {
                const uint32_t c = s_wave[k][d];
                s_wave[k][d] = cnt;
                cnt += c;
            }
        }
        uint32_t tot;
        const uint32_t start = block_exclusive_scan_t<THREADS>(cnt, &tot);
        if (dig) s_start[d] = start;
        uint32_t* mine = status + (size_t)tile * 256 + d;
        uint32_t excl = 0;
        if (tile == 0) {  // block-uniform: the first tile starts at the digit's global offset
            excl = block_exclusive_scan_t<THREADS>(dig ? totals[d] : 0u, &tot);
        } else if (dig) {
            __hip_atomic_store(mine, kLbCount | cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // look back kLbWindow tiles per round with independent loads, nearest first; stop at
            // the first inclusive prefix (tile 0 always publishes one), wait at a tile not ready
            int j = (int)tile - 1;
            uint32_t spins = 0;
            bool done = false;
            while (!done) {
                uint32_t v[kLbWindow];
#pragma unroll
                for (int k = 0; k < kLbWindow; ++k)
                    v[k] = j - k >= 0 ? __hip_atomic_load(status + (size_t)(j - k) * 256 + d,
                                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                      : kLbPrefix;
                int used = 0;
#pragma unroll
                for (int k = 0; k < kLbWindow; ++k) {
                    if (done || used < k) continue;  // stopped earlier in this round
                    if ((v[k] & ~kLbMask) == 0u) continue;  // not ready: poll again from here
                    excl += v[k] & kLbMask;
                    done = (v[k] & kLbPrefix) != 0u;
                    used = k + 1;
                }
                j -= used;
                if (!done && used < kLbWindow) {
                    // never expected: record it and stop waiting. The tile's prefix is then
                    // short of the true one, so this and every later tile still place their
                    // keys inside [0, n)
                    if (++spins > spin_limit) {
                        __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        if (dig) {
            __hip_atomic_store(mine, kLbPrefix | (excl + cnt), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            s_base[d] = excl;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        if (rank[r] == 0xFFFFFFFFu) continue;
        const uint32_t d = (uint32_t)(key[r] >> shift) & 0xFFu;
        const uint32_t pos = s_start[d] + s_wave[w][d] + rank[r];
        s_keys[pos] = key[r];
        s_vals[pos] = val[r];
    }
    __syncthreads();
    const uint32_t tile_n = min((uint32_t)kTile, n - tile * kTile);
    for (uint32_t p = threadIdx.x; p < tile_n; p += THREADS) {
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
                                                   float4* __restrict__ pos4,
                                                   const uint32_t* __restrict__ sort_err,
                                                   void* __restrict__ codes, int wide) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    // a sort that gave up leaves no valid permutation: keep the input order and make every code
    // 0 (sorted), so the tree build sees a well-formed input; the build reports the error
    const bool failed = sort_err[0] != 0u;
    const uint32_t j = failed ? i : min(perm[i], n - 1);
    if (failed) {
        if (wide)
            static_cast<uint64_t*>(codes)[i] = 0ull;
        else
            static_cast<uint32_t*>(codes)[i] = 0u;
    }
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
                                                        int B, int R, int cap, int fixed_depth,
                                                        uint8_t* __restrict__ leaf_lvl,
                                                        uint8_t* __restrict__ lcp1,
                                                        uint32_t* __restrict__ cnt,
                                                        const uint32_t* __restrict__ status) {
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
        // w(j) = common level of particles j and j+cap: the window [j, j+cap] shares a cell of
        // that level, so a cell holding i splits while some window through i shares it
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
    if (adaptive) {
        int m = -1;
        for (int t = 0; t <= cap; ++t) m = max(m, (int)s_w[threadIdx.x + t]);
        leaf = max(min(m + 1, B), R);  // more than `cap` in the finest cell: stop at depth B
    } else {
        leaf = fixed_depth;
    }
    // levels below R only separate batched systems: every system's tree starts at level R
    const int lp = max((i == 0) ? -1 : lcp_level(codes[i - 1], codes[i], B), R - 1);
    leaf_lvl[i] = (uint8_t)leaf;
    lcp1[i] = (uint8_t)(lp + 1);
    // a build whose sort gave up or whose input is not finite gets an empty tree: every walk
    // over it ends at once (the call reports the error)
    const bool failed = (status[0] | status[1]) != 0u;
    cnt[i] = (!failed && leaf > lp) ? (uint32_t)(leaf - lp) : 0u;
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

// nodes of particle i: k = noff[i] + t at level lcp1[i] + t, t < cnt[i]
__global__ __launch_bounds__(kBlock) void k_tree_emit_heads(uint32_t n,
                                                            const uint8_t* __restrict__ lcp1,
                                                            const uint32_t* __restrict__ cnt,
                                                            const uint32_t* __restrict__ noff,
                                                            Node* __restrict__ nodes,
                                                            uint4* __restrict__ child4,
                                                            uint32_t cap_nodes) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = cnt[i];
    const uint32_t base = noff[i];
    const uint32_t l0 = lcp1[i];
    for (uint32_t t = 0; t < c; ++t) {
        const uint32_t k = base + t;
        if (k >= cap_nodes) break;  // cannot happen for the capacity the context allocates
        nodes[k].first = i;
        nodes[k].meta = (l0 + t) | (t == 0 ? kMetaHead : 0u) | (t + 1 == c ? kMetaLeaf : 0u);
        child4[k] = make_uint4(kNone, kNone, kNone, kNone);
    }
}

// one thread per node: subtree extent, flags, side, parent and the parent's child slot
template <typename K>
__global__ __launch_bounds__(kBlock) void k_tree_emit_nodes(const K* __restrict__ codes,
                                                            uint32_t n, int B, int R,
                                                            const uint8_t* __restrict__ lcp1,
                                                            const uint32_t* __restrict__ noff,
                                                            const BuildParams* __restrict__ prm,
                                                            const BuildParams* __restrict__ sys_prm,
                                                            int single_child_pp,
                                                            Node* __restrict__ nodes,
                                                            uint32_t* __restrict__ parent,
                                                            uint32_t* __restrict__ child4,
                                                            uint32_t cap_nodes) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t total = prm->n_nodes;
    if (k >= total || k >= cap_nodes) return;
    const uint32_t i = nodes[k].first;
    uint32_t meta = nodes[k].meta;
    const int l = (int)(meta & kMetaLevelMask);
    const K ci = codes[i];
    uint32_t e;
    if (l == 0) {  // the single root of a one-system tree (R = 0)
        e = n;
    } else {
        const int sh = 2 * (B - l);
        const K hk = ci | (K)(((K)1 << sh) - 1);
        e = upper_from(codes, n, hk, i + 1);
    }
    if (meta & kMetaLeaf) {
        meta |= kMetaPP;
    } else if (single_child_pp) {
        const int shc = 2 * (B - l - 1);
        if (((uint32_t)(ci >> shc) & 3u) == ((uint32_t)(codes[e - 1] >> shc) & 3u)) meta |= kMetaPP;
    }
    const BuildParams* sp = R ? sys_prm + (uint32_t)(ci >> (2 * (B - R))) : prm;
    const float sl = ldexpf(sp->side, -(l - R));
    Node* nd = nodes + k;
    nd->l2 = sl * sl;
    nd->skip = (e < n) ? noff[e] : total;
    nd->count = e - i;
    nd->meta = meta;
    uint32_t par = kNone;
    if (l > R) {
        if (!(meta & kMetaHead)) {
            par = k - 1;
        } else {
            const int pl = l - 1;
            uint32_t ip = 0;
            if (pl > 0) {
                const int psh = 2 * (B - pl);
                const K lk = ci & ~(K)(((K)1 << psh) - 1);
                ip = lower_back(codes, lk, i);
            }
            par = noff[ip] + (uint32_t)(pl - (int)lcp1[ip]);
        }
        child4[4ull * par + ((uint32_t)(ci >> (2 * (B - l))) & 3u)] = k;
    }
    parent[k] = par;
}

// ------------------------------------------------------------------------------------------
// upward pass: mass, centre of mass and (optionally) second moments
//
// Pass 1: each block owns a chunk of kUpChunk consecutive nodes and settles them level by
// level (deepest first) in LDS; a node whose subtree runs past the chunk end is left for pass
// 2. Those nodes are ancestors of the chunk's end, so there are at most kMaxLevels + 1 of them
// per chunk. Pass 2: one block settles all of them level by level.
// Children are combined in digit order, so the result does not depend on scheduling.
// ------------------------------------------------------------------------------------------
struct Moments {
    float m, x, y, xx, xy, yy;
};

__device__ __forceinline__ Moments leaf_moments(const float4* __restrict__ pos4, uint32_t f,
                                                uint32_t c, bool with_m2) {
    const float4 p0 = pos4[f];
    float M = 0.f, sx = 0.f, sy = 0.f;
    // kLeafBatch loads in flight at a time (pos4 is padded past n; the tail is masked), summed
    // in particle order
    constexpr uint32_t kLeafBatch = 8;
    for (uint32_t j0 = f; j0 < f + c; j0 += kLeafBatch) {
        float4 q[kLeafBatch];
#pragma unroll
        for (uint32_t u = 0; u < kLeafBatch; ++u) q[u] = pos4[min(j0 + u, f + c - 1)];
#pragma unroll
        for (uint32_t u = 0; u < kLeafBatch; ++u) {
            if (j0 + u >= f + c) break;
            M += q[u].z;
            sx = fmaf(q[u].z, q[u].x - p0.x, sx);
            sy = fmaf(q[u].z, q[u].y - p0.y, sy);
        }
    }
    const float inv = (M != 0.f) ? 1.0f / M : 0.f;
    Moments r{M, p0.x + sx * inv, p0.y + sy * inv, 0.f, 0.f, 0.f};
    if (with_m2) {
        for (uint32_t j = f; j < f + c; ++j) {
            const float4 p = pos4[j];
            const float rx = p.x - r.x, ry = p.y - r.y;
            r.xx = fmaf(p.z * rx, rx, r.xx);
            r.xy = fmaf(p.z * rx, ry, r.xy);
            r.yy = fmaf(p.z * ry, ry, r.yy);
        }
    }
    return r;
}

// Fixed-depth trees can hold thousands of points in one leaf (a tight cluster): such leaves
// (more than kBigLeaf points) get their mass and centre of mass from one wave each before the
// chunked pass, which then takes them from the node record.
constexpr uint32_t kBigLeaf = 64;

__global__ __launch_bounds__(kBlock) void k_big_leaf_moments(const float4* __restrict__ pos4,
                                                             Node* __restrict__ nodes,
                                                             const BuildParams* __restrict__ prm,
                                                             uint32_t cap_nodes) {
    const int lane = threadIdx.x & 63;
    const uint32_t k = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (k >= cap_nodes || k >= prm->n_nodes) return;
    const uint32_t meta = nodes[k].meta, f = nodes[k].first, c = nodes[k].count;
    if (!(meta & kMetaLeaf) || c <= kBigLeaf) return;
    const float4 p0 = pos4[f];
    float M = 0.f, sx = 0.f, sy = 0.f;
    for (uint32_t j = f + lane; j < f + c; j += kWave) {
        const float4 q = pos4[j];
        M += q.z;
        sx = fmaf(q.z, q.x - p0.x, sx);
        sy = fmaf(q.z, q.y - p0.y, sy);
    }
    M = wave_sum_f(M);
    sx = wave_sum_f(sx);
    sy = wave_sum_f(sy);
    if (lane == 0) {
        const float inv = (M != 0.f) ? 1.0f / M : 0.f;
        nodes[k].cx = p0.x + sx * inv;
        nodes[k].cy = p0.y + sy * inv;
        nodes[k].mass = M;
    }
}

// combine up to four children (absent ones have m = 0 and are flagged by has[t] = false)
__device__ __forceinline__ Moments combine(const Moments* ch, const bool* has, bool with_m2) {
    // origin: the first child present (selected with constant indices: no private memory)
    float ox = ch[3].x, oy = ch[3].y;
#pragma unroll
    for (int t = 2; t >= 0; --t) {
        if (has[t]) {
            ox = ch[t].x;
            oy = ch[t].y;
        }
    }
    float M = 0.f, px = 0.f, py = 0.f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        if (!has[t]) continue;
        M += ch[t].m;
        px = fmaf(ch[t].m, ch[t].x - ox, px);
        py = fmaf(ch[t].m, ch[t].y - oy, py);
    }
    const float inv = (M != 0.f) ? 1.0f / M : 0.f;
    Moments r{M, ox + px * inv, oy + py * inv, 0.f, 0.f, 0.f};
    if (with_m2) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            if (!has[t]) continue;
            const float rx = ch[t].x - r.x, ry = ch[t].y - r.y;
            r.xx += ch[t].xx + ch[t].m * rx * rx;
            r.xy += ch[t].xy + ch[t].m * rx * ry;
            r.yy += ch[t].yy + ch[t].m * ry * ry;
        }
    }
    return r;
}

__global__ __launch_bounds__(kUpThreads) void k_up_chunk(const float4* __restrict__ pos4,
                                                     Node* __restrict__ nodes,
                                                     const uint4* __restrict__ child4,
                                                     float4* __restrict__ m2,
                                                     const BuildParams* __restrict__ prm,
                                                     uint32_t* __restrict__ deferred,
                                                     uint32_t cap_nodes, int big_done) {
    __shared__ float s_m[kUpChunk], s_x[kUpChunk], s_y[kUpChunk];
    __shared__ float s_xx[kUpChunk], s_xy[kUpChunk], s_yy[kUpChunk];
    __shared__ uint32_t s_maxl;
    __shared__ uint32_t s_ndef;
    constexpr int kPer = kUpChunk / kUpThreads;
    const uint32_t total = min(prm->n_nodes, cap_nodes);
    const uint32_t c0 = blockIdx.x * kUpChunk;
    uint32_t* my_def = deferred + (size_t)blockIdx.x * kDeferStride;
    if (c0 >= total) {
        if (threadIdx.x == 0) my_def[0] = 0;
        return;
    }
    const uint32_t c1 = min(c0 + kUpChunk, total);
    const bool with_m2 = m2 != nullptr;
    if (threadIdx.x == 0) {
        s_maxl = 0;
        s_ndef = 0;
    }
    __syncthreads();
    Node nd[kPer];
    uint32_t lmax = 0;
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const uint32_t k = c0 + threadIdx.x + r * kUpThreads;
        if (k < c1) {
            nd[r] = nodes[k];
            lmax = max(lmax, nd[r].meta & kMetaLevelMask);
        }
    }
    atomicMax(&s_maxl, lmax);
    // leaves first, all at once: their particle loads overlap instead of waiting level by level
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const uint32_t k = c0 + threadIdx.x + r * kUpThreads;
        if (k >= c1 || !(nd[r].meta & kMetaLeaf)) continue;
        const uint32_t s = k - c0;
        // leaves of more than kBigLeaf points hold their moments already when k_big_leaf_moments
        // ran (fixed-depth trees); otherwise (coincident points at the depth limit of an
        // adaptive tree, rare) this thread sums them
        const Moments mo = (big_done && nd[r].count > kBigLeaf)
                               ? Moments{nd[r].mass, nd[r].cx, nd[r].cy, 0.f, 0.f, 0.f}
                               : leaf_moments(pos4, nd[r].first, nd[r].count, with_m2);
        s_m[s] = mo.m; s_x[s] = mo.x; s_y[s] = mo.y;
        s_xx[s] = mo.xx; s_xy[s] = mo.xy; s_yy[s] = mo.yy;
    }
    __syncthreads();
    for (int L = (int)s_maxl; L >= 0; --L) {
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const uint32_t k = c0 + threadIdx.x + r * kUpThreads;
            if (k >= c1 || (int)(nd[r].meta & kMetaLevelMask) != L) continue;
            const uint32_t s = k - c0;
            if (nd[r].meta & kMetaLeaf) {
                continue;
            } else if (nd[r].skip <= c1) {
                const uint4 cs = child4[k];
                const uint32_t ci[4] = {cs.x, cs.y, cs.z, cs.w};
                Moments ch[4];
                bool has[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    has[t] = ci[t] != kNone;
                    const uint32_t q = has[t] ? ci[t] - c0 : s;
                    ch[t] = Moments{s_m[q], s_x[q], s_y[q], s_xx[q], s_xy[q], s_yy[q]};
                }
                const Moments mo = combine(ch, has, with_m2);
                s_m[s] = mo.m; s_x[s] = mo.x; s_y[s] = mo.y;
                s_xx[s] = mo.xx; s_xy[s] = mo.xy; s_yy[s] = mo.yy;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const uint32_t k = c0 + threadIdx.x + r * kUpThreads;
        if (k >= c1) continue;
        const uint32_t s = k - c0;
        if ((nd[r].meta & kMetaLeaf) || nd[r].skip <= c1) {
            nodes[k].cx = s_x[s];
            nodes[k].cy = s_y[s];
            nodes[k].mass = s_m[s];
            if (with_m2) m2[k] = make_float4(s_xx[s], s_xy[s], s_yy[s], 0.f);
        } else {
            // ancestors of the chunk end (at most one per level)
            const uint32_t slot = atomicAdd(&s_ndef, 1u);
            if (slot < kDeferStride - 1) my_def[1 + slot] = k;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) my_def[0] = min(s_ndef, (uint32_t)(kDeferStride - 1));
}

__device__ __forceinline__ Moments load_moments(const Node* __restrict__ nodes,
                                                const float4* __restrict__ m2, uint32_t q,
                                                bool with_m2) {
    Moments r;
    r.m = __hip_atomic_load(&nodes[q].mass, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    r.x = __hip_atomic_load(&nodes[q].cx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    r.y = __hip_atomic_load(&nodes[q].cy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    r.xx = r.xy = r.yy = 0.f;
    if (with_m2) {
        r.xx = __hip_atomic_load(&m2[q].x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        r.xy = __hip_atomic_load(&m2[q].y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        r.yy = __hip_atomic_load(&m2[q].z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    return r;
}

// Single block of kUpThreads: the nodes left over by k_up_chunk (ancestors of chunk ends),
// deepest level first. With at most kTopMax of them, everything stays on chip: the list is
// compacted into LDS, an LDS hash maps node index -> list entry, the moments of children that
// k_up_chunk settled are loaded once into registers (thread e % kUpThreads owns entry e), and
// the moments of deferred children are passed through LDS from one level to the next. Larger
// lists take the same steps through global memory. Children are combined in digit order
// either way, so the result does not depend on which nodes were deferred.
constexpr int kTopHash = 4096;                                 // hash slots, > kTopMax
constexpr int kTopPer = (kTopMax + kUpThreads - 1) / kUpThreads;  // entries per thread
static_assert(kTopHash >= kTopMax && (kTopHash & (kTopHash - 1)) == 0, "hash size");
static_assert(kTopMax < 65536, "entry indices are 16-bit");

__device__ __forceinline__ uint32_t top_hash(uint32_t k) { return (k * 0x9E3779B1u) >> 20; }

__global__ __launch_bounds__(kUpThreads) void k_up_top(Node* __restrict__ nodes,
                                                       const uint4* __restrict__ child4,
                                                       float4* __restrict__ m2,
                                                       const BuildParams* __restrict__ prm,
                                                       const uint32_t* __restrict__ deferred,
                                                       uint32_t* __restrict__ bylevel,
                                                       uint32_t nchunks, uint32_t cap_nodes) {
    __shared__ uint32_t s_k[kTopMax];
    __shared__ uint8_t s_l[kTopMax];
    __shared__ uint32_t s_hk[kTopHash];  // node index + 1, 0 = empty
    __shared__ uint16_t s_hv[kTopHash];  // its entry
    __shared__ float s_mo[6][kTopMax];   // moments of settled entries
    __shared__ uint32_t s_count, s_maxl;
    const uint32_t tid = threadIdx.x;
    const uint32_t total = min(prm->n_nodes, cap_nodes);
    const uint32_t nch = min(nchunks, (total + kUpChunk - 1) / kUpChunk);
    const bool with_m2 = m2 != nullptr;
    if (tid == 0) {
        s_count = 0;
        s_maxl = 0;
    }
    for (uint32_t i = tid; i < kTopHash; i += kUpThreads) s_hk[i] = 0u;
    __syncthreads();
    // compaction: chunk lists in chunk order
    for (uint32_t b0 = 0; b0 < nch; b0 += kUpThreads) {
        const uint32_t b = b0 + tid;
        const uint32_t c = (b < nch) ? deferred[(size_t)b * kDeferStride] : 0u;
        uint32_t tot;
        const uint32_t off = block_exclusive_scan_t<kUpThreads>(c, &tot);
        const uint32_t base = s_count;
        for (uint32_t j = 0; j < c; ++j) {
            const uint32_t at = base + off + j;
            if (at < kTopMax) s_k[at] = deferred[(size_t)b * kDeferStride + 1 + j];
        }
        __syncthreads();
        if (tid == 0) s_count = base + tot;
        __syncthreads();
    }
    const uint32_t D = s_count;
    if (D > kTopMax) {
        // large list (many trees: batched systems): the entries bucketed by level with a
        // counting sort into `bylevel`, then each level's bucket settled, deepest first, through
        // global memory. Each node combines its settled children in digit order, so the order
        // inside a bucket does not change any result.
        __shared__ uint32_t s_lc[256], s_lo[256];  // per level: entries, then bucket start
        for (uint32_t i = tid; i < 256; i += kUpThreads) s_lc[i] = 0u;
        __syncthreads();
        const uint32_t entries = nch * (kDeferStride - 1);
        for (uint32_t e = tid; e < entries; e += kUpThreads) {
            const uint32_t b = e / (kDeferStride - 1), slot = e % (kDeferStride - 1);
            const uint32_t* lst = deferred + (size_t)b * kDeferStride;
            if (slot < lst[0]) atomicAdd(&s_lc[nodes[lst[1 + slot]].meta & kMetaLevelMask], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t at = 0;
            for (int l = 0; l < 256; ++l) {
                s_lo[l] = at;
                at += s_lc[l];
                s_lc[l] = 0u;
            }
        }
        __syncthreads();
        for (uint32_t e = tid; e < entries; e += kUpThreads) {
            const uint32_t b = e / (kDeferStride - 1), slot = e % (kDeferStride - 1);
            const uint32_t* lst = deferred + (size_t)b * kDeferStride;
            if (slot >= lst[0]) continue;
            const uint32_t k = lst[1 + slot];
            const uint32_t l = nodes[k].meta & kMetaLevelMask;
            bylevel[s_lo[l] + atomicAdd(&s_lc[l], 1u)] = k;
        }
        __syncthreads();
        for (int L = 255; L >= 0; --L) {
            const uint32_t lo = s_lo[L], cnt = s_lc[L];  // the same in every thread
            if (cnt == 0u) continue;
            for (uint32_t i = tid; i < cnt; i += kUpThreads) {
                const uint32_t k = bylevel[lo + i];
                const uint4 cs = child4[k];
                const uint32_t ci[4] = {cs.x, cs.y, cs.z, cs.w};
                Moments ch[4];
                bool has[4];
#pragma unroll
                for (int t2 = 0; t2 < 4; ++t2) {
                    has[t2] = ci[t2] != kNone;
                    ch[t2] = load_moments(nodes, m2, has[t2] ? ci[t2] : k, with_m2);
                }
                const Moments mo = combine(ch, has, with_m2);
                __hip_atomic_store(&nodes[k].cx, mo.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_store(&nodes[k].cy, mo.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_store(&nodes[k].mass, mo.m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (with_m2) {
                    __hip_atomic_store(&m2[k].x, mo.xx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    __hip_atomic_store(&m2[k].y, mo.xy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    __hip_atomic_store(&m2[k].z, mo.yy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
            __syncthreads();
        }
        return;
    }
    // on chip: levels and the hash of this thread's entries
    uint32_t lmax = 0;
    uint32_t kk[kTopPer];
#pragma unroll
    for (int r = 0; r < kTopPer; ++r) {
        const uint32_t e = tid + r * kUpThreads;
        kk[r] = e < D ? s_k[e] : 0u;
        if (e >= D) continue;
        const uint32_t l = nodes[kk[r]].meta & kMetaLevelMask;
        s_l[e] = (uint8_t)l;
        lmax = max(lmax, l);
        uint32_t h = top_hash(kk[r]) & (kTopHash - 1);
        while (atomicCAS(&s_hk[h], 0u, kk[r] + 1u) != 0u) h = (h + 1) & (kTopHash - 1);
        s_hv[h] = (uint16_t)e;
    }
    atomicMax(&s_maxl, lmax);
    __syncthreads();
    // children: deferred ones by entry, settled ones loaded once
    Moments ch[kTopPer][4];
    int ce[kTopPer][4];
    bool has[kTopPer][4];
#pragma unroll
    for (int r = 0; r < kTopPer; ++r) {
        const uint32_t e = tid + r * kUpThreads;
        const uint4 cs = e < D ? child4[kk[r]] : make_uint4(kNone, kNone, kNone, kNone);
        const uint32_t ci[4] = {cs.x, cs.y, cs.z, cs.w};
#pragma unroll
        for (int t2 = 0; t2 < 4; ++t2) {
            has[r][t2] = ci[t2] != kNone;
            ce[r][t2] = -1;
            ch[r][t2] = Moments{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            if (!has[r][t2]) continue;
            uint32_t h = top_hash(ci[t2]) & (kTopHash - 1);
            uint32_t key;
            while ((key = s_hk[h]) != 0u && key != ci[t2] + 1u) h = (h + 1) & (kTopHash - 1);
            if (key != 0u) {
                ce[r][t2] = s_hv[h];
            } else {
                const Node q = nodes[ci[t2]];
                ch[r][t2].m = q.mass;
                ch[r][t2].x = q.cx;
                ch[r][t2].y = q.cy;
                if (with_m2) {
                    const float4 q2 = m2[ci[t2]];
                    ch[r][t2].xx = q2.x;
                    ch[r][t2].xy = q2.y;
                    ch[r][t2].yy = q2.z;
                }
            }
        }
    }
    for (int L = (int)s_maxl; L >= 0; --L) {
#pragma unroll
        for (int r = 0; r < kTopPer; ++r) {
            const uint32_t e = tid + r * kUpThreads;
            if (e >= D || (int)s_l[e] != L) continue;
#pragma unroll
            for (int t2 = 0; t2 < 4; ++t2) {
                const int q = ce[r][t2];
                if (q < 0) continue;
                ch[r][t2] = Moments{s_mo[0][q], s_mo[1][q], s_mo[2][q],
                                    s_mo[3][q], s_mo[4][q], s_mo[5][q]};
            }
            const Moments mo = combine(ch[r], has[r], with_m2);
            s_mo[0][e] = mo.m; s_mo[1][e] = mo.x; s_mo[2][e] = mo.y;
            s_mo[3][e] = mo.xx; s_mo[4][e] = mo.xy; s_mo[5][e] = mo.yy;
            nodes[kk[r]].cx = mo.x;
            nodes[kk[r]].cy = mo.y;
            nodes[kk[r]].mass = mo.m;
            if (with_m2) m2[kk[r]] = make_float4(mo.xx, mo.xy, mo.yy, 0.f);
        }
        __syncthreads();
    }
}

// batched systems: the root of system s is the first node of its first particle
__global__ __launch_bounds__(kBlock) void k_batch_roots(const uint32_t* __restrict__ sys_off,
                                                        uint32_t nsys,
                                                        const uint32_t* __restrict__ noff,
                                                        const BuildParams* __restrict__ prm,
                                                        uint32_t* __restrict__ roots) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= nsys) return;
    const uint32_t r = sys_off[s] < sys_off[s + 1] ? noff[sys_off[s]] : kNone;
    roots[s] = r < prm->n_nodes ? r : kNone;  // an empty tree (failed build) has no roots
}

// tree statistics on request: deepest leaf level and leaves at the depth limit holding more
// than the leaf capacity (one partial per block, summed by the host)
__global__ __launch_bounds__(kBlock) void k_tree_stats(const Node* __restrict__ nodes,
                                                       const BuildParams* __restrict__ prm,
                                                       uint32_t cap_nodes, int depth_limit,
                                                       uint32_t leaf_cap,
                                                       uint2* __restrict__ out) {
    uint32_t mx = 0, over = 0;
    const uint32_t total = min(prm->n_nodes, cap_nodes);
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < total; k += gridDim.x * kBlock) {
        const uint32_t meta = nodes[k].meta;
        if (!(meta & kMetaLeaf)) continue;
        const uint32_t l = meta & kMetaLevelMask;
        mx = max(mx, l);
        if ((int)l == depth_limit && nodes[k].count > leaf_cap) ++over;
    }
    for (int o = 32; o > 0; o >>= 1) {
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, o, kWave));
        over += (uint32_t)__shfl_xor((int)over, o, kWave);
    }
    __shared__ uint2 red[kWavesPerBlock];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = make_uint2(mx, over);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint2 r = red[0];
        for (int w = 1; w < kWavesPerBlock; ++w) {
            r.x = max(r.x, red[w].x);
            r.y += red[w].y;
        }
        out[blockIdx.x] = r;
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

// the three slot arrays cleared in one launch
__global__ __launch_bounds__(kBlock) void k_clear_slots(uint32_t* __restrict__ a,
                                                        uint32_t* __restrict__ b,
                                                        uint32_t* __restrict__ c, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        a[i] = 0u;
        b[i] = 0u;
        c[i] = 0u;
    }
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
                   BuildParams* prm, uint32_t tree_mode, const double* caller, uint32_t* status) {
    const uint32_t g = bounds_grid(n);
    k_bounds_partial<<<g, kBlock, 0, s>>>(xs, ys, n, partials);
    k_bounds_final<<<1, kBlock, 0, s>>>(partials, g, n, prm, tree_mode, caller ? 1 : 0,
                                        caller ? caller[0] : 0.0, caller ? caller[1] : 0.0,
                                        caller ? caller[2] : 0.0, caller ? caller[3] : 0.0,
                                        status);
}

void launch_morton_gpuref(hipStream_t s, const float* xs, const float* ys, uint32_t n,
                          const BuildParams* prm, int depth, uint32_t* keys, uint32_t* vals) {
    k_morton_gpuref<<<div_up(n, kBlock), kBlock, 0, s>>>(xs, ys, n, prm, depth, keys, vals);
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

static int sort_items(uint32_t n) { return n <= kSortSmallMax ? kSortItemsSmall : kSortItemsLarge; }

uint32_t sort_blocks(uint32_t n) { return div_up(n, sort_tile(sort_items(n))); }

// counts [8][256][tiles] + look-back words [8][tiles][256] + totals [8][256] + tickets and an
// error word, for the largest tile count of any n' <= n
uint64_t sort_scratch_words(uint32_t n) {
    const uint32_t tiles = max(sort_blocks(n), div_up(min(n, kSortSmallMax), sort_tile(kSortItemsSmall)));
    return 2ull * kSortMaxPasses * 256 * tiles + kSortMaxPasses * 256 + 16;
}

void launch_radix_sort(hipStream_t s, void* keys_a, uint32_t* vals_a, void* keys_b,
                       uint32_t* vals_b, uint32_t n, int key_bits, bool wide, uint32_t* scratch,
                       uint32_t* err, uint32_t spin_limit, bool* result_in_b) {
    const uint32_t nt = sort_blocks(n);
    const int passes = (key_bits + 7) / 8;
    uint32_t* counts = scratch;
    uint32_t* status = counts + (size_t)kSortMaxPasses * 256 * nt;
    uint32_t* totals = status + (size_t)kSortMaxPasses * 256 * nt;
    uint32_t* tickets = totals + kSortMaxPasses * 256;
    const bool small = sort_items(n) == kSortItemsSmall;
#define BH_COUNTS(K, T, I)                                                                        \
    k_radix_counts<K, T, I><<<nt, T, 0, s>>>((const K*)keys_a, n, passes, nt, counts, status)
    if (wide) {
        if (small) BH_COUNTS(uint64_t, 512, 4); else BH_COUNTS(uint64_t, 1024, 4);
    } else {
        if (small) BH_COUNTS(uint32_t, 512, 4); else BH_COUNTS(uint32_t, 1024, 4);
    }
#undef BH_COUNTS
    k_radix_totals<<<div_up(passes * 256, kWavesPerBlock), kBlock, 0, s>>>(counts, nt, passes,
                                                                            totals, tickets);
    void* kin = keys_a;
    uint32_t* vin = vals_a;
    void* kout = keys_b;
    uint32_t* vout = vals_b;
    for (int p = 0; p < passes; ++p) {
        uint32_t* st = status + (size_t)p * nt * 256;
        const uint32_t* tp = totals + p * 256;
#define BH_PASS(K, T, I)                                                                          \
    k_radix_pass<K, T, I><<<nt, T, 0, s>>>((const K*)kin, vin, (K*)kout, vout, n, 8 * p, tp, st,  \
                                           tickets + p, err, spin_limit)
        static_assert(kSortItemsLarge * kBlock == 1024 * 4 && kSortItemsSmall * kBlock == 512 * 4,
                      "pass tiles match the counting tiles");
        if (wide) {
            if (small) BH_PASS(uint64_t, 512, 4); else BH_PASS(uint64_t, 1024, 4);
        } else {
            if (small) BH_PASS(uint32_t, 512, 4); else BH_PASS(uint32_t, 1024, 4);
        }
#undef BH_PASS
        void* tk = kin; kin = kout; kout = tk;
        uint32_t* tv = vin; vin = vout; vout = tv;
    }
    *result_in_b = (passes & 1) != 0;
}

void launch_gather(hipStream_t s, const uint32_t* perm, uint32_t n, const float* x, const float* y,
                   const float* m, const float* vx, const float* vy, const uint32_t* id, float* xo,
                   float* yo, float* mo, float* vxo, float* vyo, uint32_t* ido, float4* pos4,
                   const uint32_t* sort_err, void* codes, bool wide) {
    k_gather<<<div_up(n, kBlock), kBlock, 0, s>>>(perm, n, x, y, m, vx, vy, id, xo, yo, mo, vxo,
                                                  vyo, ido, pos4, sort_err, codes, wide ? 1 : 0);
}

void launch_tree(hipStream_t s, const void* codes, bool wide, uint32_t n, int B, int cap,
                 int fixed_depth, int single_child_pp, uint8_t* leaf_lvl, uint8_t* lcp1,
                 uint32_t* cnt, uint32_t* noff, uint32_t* scan_partials, BuildParams* prm,
                 Node* nodes, uint32_t* parent, uint32_t* child4, uint32_t cap_nodes, int R,
                 const BuildParams* sys_prm, const uint32_t* status) {
    const uint32_t g = div_up(n, kBlock);
    if (wide)
        k_tree_levels<uint64_t><<<g, kBlock, 0, s>>>((const uint64_t*)codes, n, B, R, cap,
                                                     fixed_depth, leaf_lvl, lcp1, cnt, status);
    else
        k_tree_levels<uint32_t><<<g, kBlock, 0, s>>>((const uint32_t*)codes, n, B, R, cap,
                                                     fixed_depth, leaf_lvl, lcp1, cnt, status);
    launch_exclusive_scan(s, cnt, noff, n, scan_partials, &prm->n_nodes);
    k_tree_emit_heads<<<g, kBlock, 0, s>>>(n, lcp1, cnt, noff, nodes, (uint4*)child4, cap_nodes);
    const uint32_t gn = div_up(cap_nodes, kBlock);
    if (wide)
        k_tree_emit_nodes<uint64_t><<<gn, kBlock, 0, s>>>((const uint64_t*)codes, n, B, R, lcp1,
                                                          noff, prm, sys_prm, single_child_pp,
                                                          nodes, parent, child4, cap_nodes);
    else
        k_tree_emit_nodes<uint32_t><<<gn, kBlock, 0, s>>>((const uint32_t*)codes, n, B, R, lcp1,
                                                          noff, prm, sys_prm, single_child_pp,
                                                          nodes, parent, child4, cap_nodes);
}

void launch_bounds_batched(hipStream_t s, const float* xs, const float* ys, const uint2* blk,
                           uint32_t nblk, const uint32_t* blk_off, const uint32_t* sys_off,
                           uint32_t nsys, float4* partials, BuildParams* sys_prm,
                           BuildParams* prm, uint32_t* status) {
    k_bounds_seg_partial<<<nblk, kBlock, 0, s>>>(xs, ys, blk, partials);
    k_bounds_seg_final<<<nsys, kBlock, 0, s>>>(partials, blk_off, sys_off, sys_prm, prm, status);
}

void launch_morton_batched(hipStream_t s, const float* xs, const float* ys, uint32_t n,
                           const uint32_t* sys_off, uint32_t nsys, const BuildParams* sys_prm,
                           int B, uint64_t* keys, uint32_t* vals) {
    k_morton_seg<<<div_up(n, kBlock), kBlock, 0, s>>>(xs, ys, n, sys_off, nsys, sys_prm, B, keys,
                                                     vals);
}

void launch_batch_roots(hipStream_t s, const uint32_t* sys_off, uint32_t nsys,
                        const uint32_t* noff, const BuildParams* prm, uint32_t* roots) {
    k_batch_roots<<<div_up(nsys, kBlock), kBlock, 0, s>>>(sys_off, nsys, noff, prm, roots);
}

uint32_t up_chunks(uint32_t cap_nodes) { return div_up(cap_nodes, kUpChunk); }

void launch_upward(hipStream_t s, const float4* pos4, Node* nodes, const uint32_t* child4,
                   float4* m2, const BuildParams* prm, uint32_t* deferred, uint32_t* bylevel,
                   uint32_t cap_nodes, bool big_leaves) {
    const uint32_t nch = up_chunks(cap_nodes);
    const bool big_done = big_leaves && !m2;
    if (big_done)
        k_big_leaf_moments<<<div_up(cap_nodes, kWavesPerBlock), kBlock, 0, s>>>(pos4, nodes, prm,
                                                                                cap_nodes);
    k_up_chunk<<<nch, kUpThreads, 0, s>>>(pos4, nodes, (const uint4*)child4, m2, prm, deferred,
                                      cap_nodes, big_done ? 1 : 0);
    k_up_top<<<1, kUpThreads, 0, s>>>(nodes, (const uint4*)child4, m2, prm, deferred, bylevel,
                                      nch, cap_nodes);
}

void launch_tree_stats(hipStream_t s, const Node* nodes, const BuildParams* prm,
                       uint32_t cap_nodes, int depth_limit, uint32_t leaf_cap, uint2* out,
                       uint32_t blocks) {
    k_tree_stats<<<blocks, kBlock, 0, s>>>(nodes, prm, cap_nodes, depth_limit, leaf_cap, out);
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
    k_clear_slots<<<div_up(slots, kBlock), kBlock, 0, s>>>(first_child, first_particle, count,
                                                           slots);
    k_tree_slots<<<div_up(cap_nodes, kBlock), kBlock, 0, s>>>(codes, nodes, prm, cap_nodes, depth,
                                                              first_child, first_particle, count);
}

}  // namespace bh
