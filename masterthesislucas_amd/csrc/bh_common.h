// Shared device-side types and helpers of the Barnes-Hut force step (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bh_softfp.h"

namespace bh {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kMaxLevels = 32;  // deepest tree level (32 bits per axis)
constexpr int kStackDepth = 40; // > kMaxLevels + 1 pushes per traversal
constexpr int kFoldEvery = 8;   // node visits between folds of the monopole batch
constexpr int kPPUnroll = 8;    // sources per particle-particle fetch
constexpr int kPosPad = 8;      // float4 entries readable past the last particle (>= kPPUnroll)

enum : uint32_t {
    kMetaLevelMask = 0xFFu,
    kMetaLeaf = 1u << 8,   // deepest cell on its path (holds <= leaf capacity particles)
    kMetaPP = 1u << 9,     // near field is particle-particle over the whole cell
    kMetaHead = 1u << 10,  // first node starting at its particle (its parent starts earlier)
};

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kUpChunk = 2048;                 // nodes per block in the upward pass
constexpr int kUpThreads = 1024;               // threads per block in the upward pass
constexpr int kDeferStride = kMaxLevels + 2;   // count + at most one deferred node per level
constexpr int kTopMax = 2560;                  // deferred nodes the top pass keeps in LDS

enum : uint32_t {
    kTreeQuadTree = 0,     // adaptive, leaf capacity, QuadTree root and near-field rules
    kTreeMortonFixed = 1,  // fixed depth, reference 32-bit Morton codes
    kTreeGpuRef = 2,       // fixed depth, centred root, split at the cell centre (<= goes low)
};

// Per-build parameters, produced on the device by the bounds pass and the tree build.
struct BuildParams {
    float minx, miny, maxx, maxy;  // bounding box (FP32, exact min/max of the positions)
    uint64_t minx64, miny64;       // binary64 bits of the root corner
    uint64_t neg_minx64, neg_miny64;
    uint64_t side64;               // binary64 bits of the root side S
    bh_sfp_quant quant;            // floor(2^B RN64(d / S)) quantiser
    float side;                    // RN32(S)
    uint32_t n;                    // particle count
    uint32_t n_nodes;              // nodes in the pre-order tree
    uint32_t max_level;            // deepest leaf level seen
    uint32_t overfull_leaves;      // leaves at the depth limit holding > leaf capacity
    uint32_t pad[3];
};

// Traversal record, 32 bytes: read once per visit by the whole wave (scalar loads).
struct __attribute__((aligned(16))) Node {
    float cx, cy;   // centre of mass
    float mass;     // total mass
    float l2;       // side^2
    uint32_t skip;  // next node in pre-order after this subtree
    uint32_t first; // first particle (sorted order)
    uint32_t count; // particles in the cell
    uint32_t meta;  // level | kMetaLeaf | kMetaPP
};

__device__ __forceinline__ uint64_t lane_mask_lt() {
    const int lane = threadIdx.x & (kWave - 1);
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// Level of the longest common prefix (2 bits per level) of two codes with B bits per axis.
__device__ __forceinline__ int lcp_level(uint64_t a, uint64_t b, int B) {
    const uint64_t x = a ^ b;
    if (!x) return B;
    const int lz = __builtin_clzll(x) - (64 - 2 * B);
    return lz >> 1;
}

// blockIdx -> work block such that each XCD (blocks are dealt round-robin over 8 XCDs) gets a
// contiguous range of work blocks, which keeps neighbouring tree cells in one L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t q = nb >> 3, r = nb & 7u;
    const uint32_t x = b & 7u, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ float wave_sum_f(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ __forceinline__ double wave_sum_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

}  // namespace bh
