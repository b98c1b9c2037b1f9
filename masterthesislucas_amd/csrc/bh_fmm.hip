// Symmetric dual-tree FMM with the t-SNE kernels on the fixed-depth Morton tree, gfx950, FP64.
//
// Reference: NBodySolverFMM_SYM_MORTON (nBodySolverFMM_SYM_MORTON.h): traverse_SYM_NN
// (:154-270) recurses over pairs of nodes from (root, root); a pair of nodes whose distance
// passes (l_A + l_B)/theta interacts through expansions (TSNE_FMM_SYM_MORTON_NN_Kernel,
// :622-723), a pair of leaves that does not is resolved point by point: each point of the
// first leaf either interacts with the second leaf as a whole (point-node kernel, :726-816) or
// with each of its points (point-point kernel, :818-836). Both sides of every interaction are
// updated. applyForces (:482-536) then translates the local expansions down the tree and
// evaluates them at the points.
//
// GPU formulation (no pair lists, no atomics, fixed summation order):
//  - every node X walks the implicit quadtree from the root, pairing each visited node Y at
//    level k with X's ancestor at level k, and descends into Y only while that pair is not
//    accepted. The accepted pairs at X's own level are exactly the node pairs the recursion
//    gives X (all nodes at one level share one box size, and both sides of a pair are split
//    together), so X gathers the expansions of all its partners (k_fmm_nn, one thread per
//    node). The reference's symmetric update adds to the second node what this gather computes
//    when that node is the sink.
//  - every leaf X repeats the walk to its leaf partners (the non-accepted leaf pairs). The
//    reference orients a leaf pair (A, B) with A the lower slot (children are paired in slot
//    order), and only A's points are tested against B's box; the sink leaf takes both roles:
//    as A its points get the point-node or point-point sums, as B it gets, per point q of the
//    partner, the point-point terms (q not accepted) or the point-to-node expansion into its
//    own coefficients (q accepted) (k_fmm_leaf, one wave per leaf, lanes over its points).
//  - the downward pass runs level by level; the leaf evaluation adds the local expansion.
//  Z: every directed evaluation adds half of the reference's symmetric total.
#include "bh_common.h"
#include "bh_launch.h"

namespace bh {

__device__ __forceinline__ uint32_t fmm_level_index(int l) {
    return (uint32_t)((((uint64_t)1 << (2 * l)) - 1) / 3);
}

// Coefficients of one node, structure of arrays over the slots: C1 (2), C2 (4: [i][j] at 2i+j),
// C3 (8: [i][j][k] at 4i+2j+k)
struct FmmCoef {
    double c1[2], c2[4], c3[8];
};

__device__ __forceinline__ void fmm_derivs(double rx, double ry, double& rS, double& D1,
                                           double& D2, double& D3) {
    rS = 1.0 + (rx * rx + ry * ry);
    const double inv = 1.0 / rS, inv2 = inv * inv, inv3 = inv2 * inv, inv4 = inv3 * inv;
    D1 = inv2;
    D2 = -4.0 * inv3;
    D3 = 24.0 * inv4;
}

// C1 at offset R from a source of mass mb and second moment (xx, xy, yy)
__device__ __forceinline__ void fmm_c1(double rx, double ry, double D1, double D2, double D3,
                                       double mb, double xx, double xy, double yy, double& ox,
                                       double& oy) {
    const double t00 = xx / mb, t01 = xy / mb, t10 = xy / mb, t11 = yy / mb;
    const double s1 = t00 + t11;
    const double s2 = (rx * rx * t00) + (rx * ry * t01) + (ry * rx * t10) + (ry * ry * t11);
    const double s3x = rx * t00 + ry * t01, s3y = rx * t10 + ry * t11;
    ox = mb * (rx * (D1 + 0.5 * s1 * D2 + 0.5 * s2 * D3) + s3x * D2);
    oy = mb * (ry * (D1 + 0.5 * s1 * D2 + 0.5 * s2 * D3) + s3y * D2);
}

__device__ __forceinline__ void fmm_add_c2c3(double rx, double ry, double D1, double D2, double D3,
                                             double f, FmmCoef& c) {
    c.c2[0] += f * (D1 + rx * rx * D2);
    c.c2[1] += f * (rx * ry * D2);
    c.c2[2] += f * (ry * rx * D2);
    c.c2[3] += f * (D1 + ry * ry * D2);
    c.c3[0] += f * ((rx + rx + rx) * D2 + rx * rx * rx * D3);
    c.c3[1] += f * (ry * D2 + rx * rx * ry * D3);
    c.c3[2] += f * (ry * D2 + rx * ry * rx * D3);
    c.c3[3] += f * (rx * D2 + rx * ry * ry * D3);
    c.c3[4] += f * (ry * D2 + ry * rx * rx * D3);
    c.c3[5] += f * (rx * D2 + ry * rx * ry * D3);
    c.c3[6] += f * (rx * D2 + ry * ry * rx * D3);
    c.c3[7] += f * ((ry + ry + ry) * D2 + ry * ry * ry * D3);
}

// v = C1 + d.C2 + 0.5 d.(d.C3), t = C2 + d.C3 (contractions over the first index)
__device__ __forceinline__ void fmm_translate(const FmmCoef& c, double dx, double dy, double& vx,
                                              double& vy, double* t2) {
    double t[4];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int k = 0; k < 2; ++k) t[2 * j + k] = dx * c.c3[2 * j + k] + dy * c.c3[4 + 2 * j + k];
    vx = c.c1[0] + (dx * c.c2[0] + dy * c.c2[2]) + 0.5 * (dx * t[0] + dy * t[2]);
    vy = c.c1[1] + (dx * c.c2[1] + dy * c.c2[3]) + 0.5 * (dx * t[1] + dy * t[3]);
    if (t2) {
#pragma unroll
        for (int k = 0; k < 4; ++k) t2[k] = c.c2[k] + t[k];
    }
}

__device__ __forceinline__ FmmCoef fmm_load(const double* __restrict__ C, uint32_t S, uint32_t s) {
    FmmCoef c;
#pragma unroll
    for (int k = 0; k < 2; ++k) c.c1[k] = C[(size_t)k * S + s];
#pragma unroll
    for (int k = 0; k < 4; ++k) c.c2[k] = C[(size_t)(2 + k) * S + s];
#pragma unroll
    for (int k = 0; k < 8; ++k) c.c3[k] = C[(size_t)(6 + k) * S + s];
    return c;
}

__device__ __forceinline__ void fmm_store(double* __restrict__ C, uint32_t S, uint32_t s,
                                          const FmmCoef& c) {
#pragma unroll
    for (int k = 0; k < 2; ++k) C[(size_t)k * S + s] = c.c1[k];
#pragma unroll
    for (int k = 0; k < 4; ++k) C[(size_t)(2 + k) * S + s] = c.c2[k];
#pragma unroll
    for (int k = 0; k < 8; ++k) C[(size_t)(6 + k) * S + s] = c.c3[k];
}

// node moments, structure of arrays: m0, cx, cy, m2xx, m2xy, m2yy (S each)
struct FmmMoments {
    double *m0, *cx, *cy, *xx, *xy, *yy;
};

__device__ __forceinline__ FmmMoments fmm_moments(double* base, uint32_t S) {
    return FmmMoments{base, base + S, base + 2 * (size_t)S, base + 3 * (size_t)S,
                      base + 4 * (size_t)S, base + 5 * (size_t)S};
}

// createLeafNodes (:328-379): COM = mean of the leaf's points, M2 = sum of (p-com)(p-com)^T.
// One wave per leaf, lanes striding over its points (four loads in flight per lane). Also
// writes the FP64 copy of the positions the point-point sums read.
__global__ __launch_bounds__(kBlock) void k_fmm_leaf_moments(const float4* __restrict__ pos4,
                                                             const uint32_t* __restrict__ first,
                                                             const uint32_t* __restrict__ count,
                                                             int depth, uint32_t S,
                                                             double* __restrict__ mom,
                                                             double2* __restrict__ pd) {
    const int lane = threadIdx.x & 63;
    const uint32_t i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (i >= (1u << (2 * depth))) return;
    const uint32_t s = fmm_level_index(depth) + i;
    FmmMoments M = fmm_moments(mom, S);
    const uint32_t f = first[s], c = count[s], e = f + c;
    double xx = 0.0, xy = 0.0, yy = 0.0, m0 = 0.0, cx = 0.0, cy = 0.0;
    if (c) {
        double sx = 0.0, sy = 0.0;
        uint32_t p = f + lane;
        for (; p + 3 * kWave < e; p += 4 * kWave) {
            const float4 a0 = pos4[p], a1 = pos4[p + kWave], a2 = pos4[p + 2 * kWave],
                         a3 = pos4[p + 3 * kWave];
            sx += ((double)a0.x + (double)a1.x) + ((double)a2.x + (double)a3.x);
            sy += ((double)a0.y + (double)a1.y) + ((double)a2.y + (double)a3.y);
            pd[p] = make_double2(a0.x, a0.y);
            pd[p + kWave] = make_double2(a1.x, a1.y);
            pd[p + 2 * kWave] = make_double2(a2.x, a2.y);
            pd[p + 3 * kWave] = make_double2(a3.x, a3.y);
        }
        for (; p < e; p += kWave) {
            const float4 a0 = pos4[p];
            sx += (double)a0.x;
            sy += (double)a0.y;
            pd[p] = make_double2(a0.x, a0.y);
        }
        m0 = (double)c;
        cx = wave_sum_d(sx) / m0;
        cy = wave_sum_d(sy) / m0;
        p = f + lane;
        for (; p + 3 * kWave < e; p += 4 * kWave) {
            const float4 a0 = pos4[p], a1 = pos4[p + kWave], a2 = pos4[p + 2 * kWave],
                         a3 = pos4[p + 3 * kWave];
            const double r0x = (double)a0.x - cx, r0y = (double)a0.y - cy;
            const double r1x = (double)a1.x - cx, r1y = (double)a1.y - cy;
            const double r2x = (double)a2.x - cx, r2y = (double)a2.y - cy;
            const double r3x = (double)a3.x - cx, r3y = (double)a3.y - cy;
            xx += (r0x * r0x + r1x * r1x) + (r2x * r2x + r3x * r3x);
            xy += (r0x * r0y + r1x * r1y) + (r2x * r2y + r3x * r3y);
            yy += (r0y * r0y + r1y * r1y) + (r2y * r2y + r3y * r3y);
        }
        for (; p < e; p += kWave) {
            const float4 a0 = pos4[p];
            const double rx = (double)a0.x - cx, ry = (double)a0.y - cy;
            xx += rx * rx;
            xy += rx * ry;
            yy += ry * ry;
        }
        xx = wave_sum_d(xx);
        xy = wave_sum_d(xy);
        yy = wave_sum_d(yy);
    }
    if (lane == 0) {
        M.m0[s] = m0;
        M.cx[s] = cx;
        M.cy[s] = cy;
        M.xx[s] = xx;
        M.xy[s] = xy;
        M.yy[s] = yy;
    }
}

// bottomUpNodeConstruction (:381-470): mass-weighted COM (children taken from the last), then
// M2 = sum over children of (com_c - com)(com_c - com)^T + M2_c (not weighted by M0_c)
struct FmmMom {
    double m0, cx, cy, xx, xy, yy;
};

__device__ __forceinline__ FmmMom fmm_up_combine(const FmmMom* ch) {
    FmmMom r{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int j = 3; j >= 0; --j) {
        if (ch[j].m0 == 0.0) continue;
        r.cx += ch[j].m0 * ch[j].cx;
        r.cy += ch[j].m0 * ch[j].cy;
        r.m0 += ch[j].m0;
    }
    if (r.m0 != 0.0) {
        r.cx /= r.m0;
        r.cy /= r.m0;
        for (int j = 0; j < 4; ++j) {
            if (ch[j].m0 == 0.0) continue;
            const double rx = ch[j].cx - r.cx, ry = ch[j].cy - r.cy;
            r.xx += rx * rx;
            r.xy += rx * ry;
            r.yy += ry * ry;
            r.xx += ch[j].xx;
            r.xy += ch[j].xy;
            r.yy += ch[j].yy;
        }
    }
    return r;
}

__device__ __forceinline__ FmmMom fmm_up_node(int l, uint32_t i, uint32_t S,
                                              double* __restrict__ mom) {
    FmmMoments M = fmm_moments(mom, S);
    const uint32_t s = fmm_level_index(l) + i, pfc = fmm_level_index(l + 1) + 4 * i;
    FmmMom ch[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
        ch[j] = FmmMom{M.m0[pfc + j], M.cx[pfc + j], M.cy[pfc + j], M.xx[pfc + j], M.xy[pfc + j],
                       M.yy[pfc + j]};
    const FmmMom r = fmm_up_combine(ch);
    M.m0[s] = r.m0;
    M.cx[s] = r.cx;
    M.cy[s] = r.cy;
    M.xx[s] = r.xx;
    M.xy[s] = r.xy;
    M.yy[s] = r.yy;
    return r;
}

// levels with at most kFmmTopNodes nodes are done by one block, level by level
constexpr uint32_t kFmmTopNodes = 1024;
constexpr int kFmmTopLevel = 5;  // 4^5 = kFmmTopNodes

__global__ __launch_bounds__(kBlock) void k_fmm_up(int l, uint32_t S, double* __restrict__ mom) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < (1u << (2 * l))) (void)fmm_up_node(l, i, S, mom);
}

// levels top..0 in one block: level top from memory, the levels above it from LDS
__global__ __launch_bounds__(kFmmTopNodes) void k_fmm_up_top(int top, uint32_t S,
                                                             double* __restrict__ mom) {
    __shared__ double sm[6][kFmmTopNodes];
    const uint32_t i = threadIdx.x;
    FmmMoments M = fmm_moments(mom, S);
    if (i < (1u << (2 * top))) {
        const FmmMom r = fmm_up_node(top, i, S, mom);
        sm[0][i] = r.m0;
        sm[1][i] = r.cx;
        sm[2][i] = r.cy;
        sm[3][i] = r.xx;
        sm[4][i] = r.xy;
        sm[5][i] = r.yy;
    }
    __syncthreads();
    for (int l = top - 1; l >= 0; --l) {
        const bool act = i < (1u << (2 * l));
        FmmMom r{};
        if (act) {
            FmmMom ch[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t c = 4 * i + j;
                ch[j] = FmmMom{sm[0][c], sm[1][c], sm[2][c], sm[3][c], sm[4][c], sm[5][c]};
            }
            r = fmm_up_combine(ch);
        }
        __syncthreads();
        if (act) {
            sm[0][i] = r.m0;
            sm[1][i] = r.cx;
            sm[2][i] = r.cy;
            sm[3][i] = r.xx;
            sm[4][i] = r.xy;
            sm[5][i] = r.yy;
            const uint32_t s = fmm_level_index(l) + i;
            M.m0[s] = r.m0;
            M.cx[s] = r.cx;
            M.cy[s] = r.cy;
            M.xx[s] = r.xx;
            M.xy[s] = r.xy;
            M.yy[s] = r.yy;
        }
        __syncthreads();
    }
}

// The walk of node X (level l, level-local index ix): visits every node Y paired with an
// ancestor of X by the reference recursion, in pre-order of the implicit quadtree; calls
// on_pair(k, Y, accepted) for each non-empty Y and descends only into non-accepted pairs below
// X's level.
template <typename F>
__device__ __forceinline__ void fmm_walk(const uint32_t* __restrict__ count, FmmMoments M,
                                         double side, double theta, int l, uint32_t ix, F&& on_pair) {
    int k = 0;
    uint32_t i = 0;
    while (true) {
        const uint32_t y = fmm_level_index(k) + i;
        bool descend = false;
        if (count[y] != 0u) {
            const uint32_t xa = fmm_level_index(k) + (ix >> (2 * (l - k)));
            const double rx = M.cx[xa] - M.cx[y], ry = M.cy[xa] - M.cy[y];
            const double dist = rx * rx + ry * ry;
            const double len = ldexp(side, -k);
            const double thr = (len + len) / theta;
            const bool acc = dist > thr * thr;
            on_pair(k, y, acc, rx, ry);
            descend = !acc && k < l;
        }
        if (descend) {
            ++k;
            i <<= 2;
            continue;
        }
        while (k > 0 && (i & 3u) == 3u) {
            i >>= 2;
            --k;
        }
        if (k == 0) break;
        ++i;
    }
}

// Interaction lists. For a node X at level l, N(X) is the set of non-empty nodes Y at level l
// whose pair with X the recursion reaches and does not accept; it is built top-down from the
// parent's list (the pairs the recursion reaches at level l are the children of the parent's
// non-accepted partners), and the accepted ones are X's node-node interactions. At the leaf
// level N(X) is the list of leaf partners. A list longer than kFmmListCap is marked as
// overflowed; nodes below it (and leaf work on it) fall back to walking the tree from the root,
// which visits the same pairs.
constexpr uint32_t kFmmListCap = 256;
constexpr uint32_t kFmmOverflow = 0xFFFFFFFFu;

// Leaf work is cut into units of (leaf X, chunk of 64 of its points, source share f of F_X):
// unit f takes every F_X-th point of each partner leaf, so one crowded leaf (a tight cluster of
// the embedding in a fixed-depth leaf) is spread over many waves. F_X grows with the number of
// points X's chunks meet (its leaf partners, itself included).
constexpr uint32_t kFmmMaxSplit = 16;
constexpr uint32_t kFmmUnitSources = 448;   // partner points per unit aimed at
constexpr uint32_t kFmmStage = 512;         // sources a wave stages in LDS (>= kFmmUnitSources + 64)
constexpr uint32_t kFmmExpSources = 2048;   // lower-partner points per expansion unit aimed at
constexpr uint32_t kFmmScanThreads = 1024;
constexpr uint32_t kFmmUnitBlocks = 4096;   // blocks of the persistent leaf-unit kernel

__device__ __forceinline__ void fmm_nn_gather(double rx, double ry, double ma, FmmMoments M,
                                              uint32_t y, FmmCoef& c, double& t) {
    double rS, D1, D2, D3;
    fmm_derivs(rx, ry, rS, D1, D2, D3);
    double vx, vy;
    fmm_c1(rx, ry, D1, D2, D3, M.m0[y], M.xx[y], M.xy[y], M.yy[y], vx, vy);
    c.c1[0] += vx;
    c.c1[1] += vy;
    fmm_add_c2c3(rx, ry, D1, D2, D3, ma * M.m0[y], c);
    t += (ma * M.m0[y]) / rS;
}

// One level of the list construction, one wave per node X of level l: candidates (children of
// the parent's partners) over the lanes; accepted ones add their expansion to X (node-node
// kernel, gathered: X is the sink), the others are appended to N(X) in candidate order. Leaf
// slots also size their work units.
__global__ __launch_bounds__(kBlock) void k_fmm_level(const uint32_t* __restrict__ count, int l,
                                                      int depth, uint32_t S,
                                                      const BuildParams* __restrict__ prm,
                                                      double theta, double* __restrict__ mom,
                                                      double* __restrict__ C,
                                                      double* __restrict__ tot,
                                                      uint32_t* __restrict__ lists,
                                                      uint32_t* __restrict__ nlist,
                                                      uint32_t* __restrict__ units,
                                                      uint32_t unit_stride,
                                                      uint32_t* __restrict__ split) {
    const int lane = threadIdx.x & 63;
    const uint32_t ix = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (ix >= (1u << (2 * l))) return;
    const uint32_t s = fmm_level_index(l) + ix;
    if (count[s] == 0u) {
        if (lane == 0) {
            nlist[s] = 0;
            tot[s] = 0.0;
            if (l == depth) {
                units[ix] = 0;
                units[unit_stride + ix] = 0;
                units[2 * unit_stride + ix] = 0;
                split[ix] = 1;
            }
        }
        return;
    }
    FmmMoments M = fmm_moments(mom, S);
    double side = __longlong_as_double((long long)prm->side64);
    if (!(side > 0.0)) side = 1.0;
    const double len = ldexp(side, -l);
    const double thr = (len + len) / theta;
    const double cxX = M.cx[s], cyX = M.cy[s], ma = M.m0[s];
    uint32_t* out = lists + (size_t)s * kFmmListCap;
    FmmCoef c = {};
    double t = 0.0;
    uint32_t meet = 0, lower = 0, n_out = 0;  // lower: points of partners of lower slot
    const uint32_t par = l > 0 ? fmm_level_index(l - 1) + (ix >> 2) : 0u;
    const uint32_t np = l > 0 ? nlist[par] : 1u;
    if (np != kFmmOverflow) {
        const uint32_t* in = lists + (size_t)par * kFmmListCap;
        const uint32_t cand = l > 0 ? 4u * np : 1u;
        const uint32_t child0 = l > 0 ? fmm_level_index(l) - 4u * fmm_level_index(l - 1) : 0u;
        for (uint32_t j0 = 0; j0 < cand; j0 += kWave) {
            const uint32_t j = j0 + lane;
            uint32_t y = 0;
            bool keep = false;
            if (j < cand) {
                y = l > 0 ? child0 + 4u * in[j >> 2] + (j & 3u) : 0u;
                if (count[y] != 0u) {
                    const double rx = cxX - M.cx[y], ry = cyX - M.cy[y];
                    if (rx * rx + ry * ry > thr * thr) fmm_nn_gather(rx, ry, ma, M, y, c, t);
                    else keep = true;
                }
            }
            const uint64_t kb = __ballot(keep);
            if (keep) {
                const uint32_t at = n_out + (uint32_t)__popcll(kb & lane_mask_lt());
                if (at < kFmmListCap) out[at] = y;
                meet += count[y];
                if (y < s) lower += count[y];
            }
            n_out += (uint32_t)__popcll(kb);
        }
        meet = (uint32_t)wave_sum_u64(meet);
        lower = (uint32_t)wave_sum_u64(lower);
    } else {
        // the parent's list overflowed: walk from the root (uniform), lanes take turns with the
        // accepted pairs
        uint32_t turn = 0;
        fmm_walk(count, M, side, theta, l, ix, [&](int k, uint32_t y, bool acc, double rx, double ry) {
            if (k != l) return;
            if (acc) {
                if ((turn++ & 63u) == (uint32_t)lane) fmm_nn_gather(rx, ry, ma, M, y, c, t);
                return;
            }
            if (n_out < kFmmListCap && lane == 0) out[n_out] = y;
            ++n_out;
            meet += count[y];
            if (y < s) lower += count[y];
        });
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) c.c1[k] = wave_sum_d(c.c1[k]);
#pragma unroll
    for (int k = 0; k < 4; ++k) c.c2[k] = wave_sum_d(c.c2[k]);
#pragma unroll
    for (int k = 0; k < 8; ++k) c.c3[k] = wave_sum_d(c.c3[k]);
    t = wave_sum_d(t);
    if (lane == 0) {
        fmm_store(C, S, s, c);
        tot[s] = t;
        nlist[s] = n_out <= kFmmListCap ? n_out : kFmmOverflow;
        if (l == depth) {
            const uint32_t chunks = (count[s] + kWave - 1) / kWave;
            uint32_t f = (meet + kFmmUnitSources - 1) / kFmmUnitSources;
            f = f < 1u ? 1u : (f > kFmmMaxSplit ? kFmmMaxSplit : f);
            // light: at most 64 partners whose shares fit the stage buffer (sum over partners of
            // ceil(n_Y / f) <= (meet + n_out f) / f)
            const bool light = n_out <= (uint32_t)kWave && meet + n_out * f <= kFmmStage * f;
            units[ix] = light ? chunks * f : 0u;
            units[unit_stride + ix] = light ? 0u : chunks * f;
            split[ix] = f;
            // expansion units: shares of the lower-slot partners' points, at most 1 + n_X/4
            // per leaf (which bounds their partial sums by 14 (leaves + n/4) doubles)
            uint32_t e = (lower + kFmmExpSources - 1) / kFmmExpSources;
            const uint32_t ecap = 1u + count[s] / 4u;
            e = e > ecap ? ecap : e;
            units[2 * unit_stride + ix] = e > kFmmMaxSplit ? kFmmMaxSplit : e;
        }
    }
}

__device__ __forceinline__ double fmm_readlane_d(double v, uint32_t k) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, (int)k);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), (int)k);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// every leaf partner Y of leaf X with its point range [fY, end) and centre: the list entries
// are fetched 64 at a time (one per lane, with their ranges and centres) and handed out with
// readlane, so stepping from one partner to the next waits on no load
template <typename Fn>
__device__ __forceinline__ void fmm_partners_meta(const uint32_t* __restrict__ first,
                                                  const uint32_t* __restrict__ count,
                                                  FmmMoments M, double side, double theta,
                                                  int depth, uint32_t X, uint32_t ix,
                                                  const uint32_t* __restrict__ lists,
                                                  const uint32_t* __restrict__ nlist, Fn&& fn) {
    const int lane = threadIdx.x & 63;
    const uint32_t np = nlist[X];
    if (np != kFmmOverflow) {
        const uint32_t* in = lists + (size_t)X * kFmmListCap;
        for (uint32_t b0 = 0; b0 < np; b0 += kWave) {
            const uint32_t j = b0 + (uint32_t)lane;
            const uint32_t yl = j < np ? in[j] : X;
            const uint32_t fl = first[yl], nl = count[yl];
            const double cxl = M.cx[yl], cyl = M.cy[yl];
            const uint32_t nb = np - b0 < (uint32_t)kWave ? np - b0 : (uint32_t)kWave;
            for (uint32_t k = 0; k < nb; ++k) {
                const uint32_t Y = (uint32_t)__builtin_amdgcn_readlane((int)yl, (int)k);
                const uint32_t fY = (uint32_t)__builtin_amdgcn_readlane((int)fl, (int)k);
                const uint32_t nY = (uint32_t)__builtin_amdgcn_readlane((int)nl, (int)k);
                fn(Y, fY, fY + nY, fmm_readlane_d(cxl, k), fmm_readlane_d(cyl, k));
            }
        }
    } else {
        fmm_walk(count, M, side, theta, depth, ix,
                 [&](int k, uint32_t y, bool acc, double, double) {
                     if (k == depth && !acc) fn(y, first[y], first[y] + count[y], M.cx[y], M.cy[y]);
                 });
    }
}

// exclusive scans of the per-leaf unit counts of the three unit classes (light, heavy,
// expansion: units + c * unit_stride, unit_stride a multiple of 4) in place; entry L of each =
// its total (one block). Each thread sums a contiguous run of entries (a multiple of 16, read
// as uint4, four loads in flight), the run totals are scanned over the block, and the runs are
// rewritten with their offsets.
constexpr int kFmmScanArrays = 3;

__device__ __forceinline__ uint4 fmm_load4(const uint32_t* __restrict__ u, uint32_t i, uint32_t L) {
    if (i + 3 < L) return *reinterpret_cast<const uint4*>(u + i);
    return make_uint4(i < L ? u[i] : 0u, i + 1 < L ? u[i + 1] : 0u, i + 2 < L ? u[i + 2] : 0u,
                      i + 3 < L ? u[i + 3] : 0u);
}

__global__ __launch_bounds__(kFmmScanThreads) void k_fmm_scan(uint32_t* __restrict__ units,
                                                              uint32_t unit_stride, uint32_t L) {
    constexpr int kW = kFmmScanThreads / kWave;
    __shared__ uint32_t wsum[kFmmScanArrays][kW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t per = ((L + kFmmScanThreads - 1) / kFmmScanThreads + 15) & ~15u;
    const uint32_t b = threadIdx.x * per;
    const uint32_t e = b + per;  // entries >= L read as 0 and are not written
    uint32_t own[kFmmScanArrays], x[kFmmScanArrays];
#pragma unroll
    for (int a = 0; a < kFmmScanArrays; ++a) {
        const uint32_t* u = units + a * unit_stride;
        uint32_t s = 0;
        for (uint32_t i = b; i < e && i < L; i += 16) {
            const uint4 v0 = fmm_load4(u, i, L), v1 = fmm_load4(u, i + 4, L),
                        v2 = fmm_load4(u, i + 8, L), v3 = fmm_load4(u, i + 12, L);
            s += (v0.x + v0.y + v0.z + v0.w) + (v1.x + v1.y + v1.z + v1.w) +
                 (v2.x + v2.y + v2.z + v2.w) + (v3.x + v3.y + v3.z + v3.w);
        }
        own[a] = s;
    }
#pragma unroll
    for (int a = 0; a < kFmmScanArrays; ++a) {
        x[a] = own[a];
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const uint32_t y = __shfl_up(x[a], d, kWave);
            if (lane >= d) x[a] += y;
        }
        if (lane == kWave - 1) wsum[a][w] = x[a];
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < kFmmScanArrays; ++a) {
        uint32_t off = 0, all = 0;
        for (int j = 0; j < kW; ++j) {
            const uint32_t s = wsum[a][j];
            off += j < w ? s : 0u;
            all += s;
        }
        uint32_t* u = units + a * unit_stride;
        uint32_t run = off + x[a] - own[a];
        for (uint32_t i = b; i < e && i < L; i += 16) {
            uint4 v[4] = {fmm_load4(u, i, L), fmm_load4(u, i + 4, L), fmm_load4(u, i + 8, L),
                          fmm_load4(u, i + 12, L)};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint4 c = v[k];
                uint4 o;
                o.x = run;
                o.y = o.x + c.x;
                o.z = o.y + c.y;
                o.w = o.z + c.z;
                run = o.w + c.w;
                const uint32_t j = i + 4 * k;
                if (j + 3 < L) {
                    *reinterpret_cast<uint4*>(u + j) = o;
                } else {
                    if (j < L) u[j] = o.x;
                    if (j + 1 < L) u[j + 1] = o.y;
                    if (j + 2 < L) u[j + 2] = o.z;
                }
            }
        }
        if (threadIdx.x == 0) u[L] = all;
    }
}

// 1 / d for d >= 1: hardware reciprocal estimate (relative error below 1e-7) refined by one
// Newton-Raphson step r <- r + r (1 - d r), which squares the error (below 1e-14)
__device__ __forceinline__ double fmm_inv(double d) {
    const double r = __builtin_amdgcn_rcp(d);
    return __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
}

// Point-to-node expansions into leaf X (the reference's point-node kernel applied to X as the
// second leaf of a pair): for every leaf partner Y of lower slot, each point q of Y farther than
// l_X/theta from X's centre adds its C1 = -R D1 and C2, C3 (x M0_X) to X. Work units (X, e of
// E_X) take the points e, e + E_X, ... of those partners over the lanes; each writes its 14
// coefficients, reduced over the wave, to epart (14 per unit), summed into X by k_fmm_exp_add.
__global__ __launch_bounds__(kBlock) void k_fmm_exp(const double2* __restrict__ pd,
                                                    const uint32_t* __restrict__ first,
                                                    const uint32_t* __restrict__ count, int depth,
                                                    uint32_t S, const BuildParams* __restrict__ prm,
                                                    double theta, double* __restrict__ mom,
                                                    const uint32_t* __restrict__ lists,
                                                    const uint32_t* __restrict__ nlist,
                                                    const uint32_t* __restrict__ unit_off,
                                                    double* __restrict__ epart) {
    const int lane = threadIdx.x & 63;
    const uint32_t L = 1u << (2 * depth);
    const uint32_t total = unit_off[L];
    const uint32_t waves = gridDim.x * kWavesPerBlock;
    FmmMoments M = fmm_moments(mom, S);
    double side = __longlong_as_double((long long)prm->side64);
    if (!(side > 0.0)) side = 1.0;
    const double thr = ldexp(side, -depth) / theta;
    const uint32_t base = fmm_level_index(depth);
    for (uint32_t u = __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
         u < total; u += waves) {
        uint32_t lo = 0, hi = L;  // unit_off[lo] <= u < unit_off[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (unit_off[mid] <= u) lo = mid;
            else hi = mid;
        }
        const uint32_t ix = lo, X = base + ix;
        const uint32_t E = unit_off[ix + 1] - unit_off[ix], e = u - unit_off[ix];
        const double cxX = M.cx[X], cyX = M.cy[X], m0X = M.m0[X];
        FmmCoef cX = {};
        const uint32_t start = e + E * (uint32_t)lane, step = E * (uint32_t)kWave;
        fmm_partners_meta(first, count, M, side, theta, depth, X, ix, lists, nlist,
                          [&](uint32_t Y, uint32_t fY, uint32_t end, double, double) {
            if (Y >= X) return;
            for (uint32_t q = fY + start; q < end; q += step) {
                const double2 qq = pd[q];
                const double rx = qq.x - cxX, ry = qq.y - cyX;
                if (rx * rx + ry * ry > thr * thr) {
                    double rS, D1, D2, D3;
                    fmm_derivs(rx, ry, rS, D1, D2, D3);
                    cX.c1[0] += -rx * D1;
                    cX.c1[1] += -ry * D1;
                    fmm_add_c2c3(-rx, -ry, D1, D2, D3, m0X, cX);
                }
            }
        });
        double v[14];
#pragma unroll
        for (int k = 0; k < 2; ++k) v[k] = wave_sum_d(cX.c1[k]);
#pragma unroll
        for (int k = 0; k < 4; ++k) v[2 + k] = wave_sum_d(cX.c2[k]);
#pragma unroll
        for (int k = 0; k < 8; ++k) v[6 + k] = wave_sum_d(cX.c3[k]);
        if (lane < 14) {
            double mine = v[0];
#pragma unroll
            for (int k = 1; k < 14; ++k) mine = lane == k ? v[k] : mine;
            epart[(size_t)u * 14 + lane] = mine;
        }
    }
}

// adds each leaf's expansion-unit partials (in unit order) to its coefficients; thread per leaf
__global__ __launch_bounds__(kBlock) void k_fmm_exp_add(const uint32_t* __restrict__ unit_off,
                                                        int depth, uint32_t S,
                                                        const double* __restrict__ epart,
                                                        double* __restrict__ C) {
    const uint32_t ix = blockIdx.x * kBlock + threadIdx.x;
    if (ix >= (1u << (2 * depth))) return;
    const uint32_t b = unit_off[ix], e = unit_off[ix + 1];
    if (b == e) return;
    const uint32_t X = fmm_level_index(depth) + ix;
    FmmCoef c = fmm_load(C, S, X);
    for (uint32_t u = b; u < e; ++u) {
        const double* p = epart + (size_t)u * 14;
#pragma unroll
        for (int k = 0; k < 2; ++k) c.c1[k] += p[k];
#pragma unroll
        for (int k = 0; k < 4; ++k) c.c2[k] += p[2 + k];
#pragma unroll
        for (int k = 0; k < 8; ++k) c.c3[k] += p[6 + k];
    }
    fmm_store(C, S, X, c);
}

// point-point term of source q on the lane's sink (p.x, p.y)
__device__ __forceinline__ void fmm_pp1(double2 q, double px, double py, double& ax, double& ay,
                                        double& t, bool count_t) {
    const double rx = px - q.x, ry = py - q.y;
    const double w = fmm_inv(__builtin_fma(rx, rx, __builtin_fma(ry, ry, 1.0)));
    const double w2 = w * w;
    ax = __builtin_fma(w2, rx, ax);
    ay = __builtin_fma(w2, ry, ay);
    t += count_t ? w : 0.0;
}

// point-point sums of one chunk against the share q, q + F, ... < end of a partner's points
// (four sources per step, their loads issued together)
__device__ __forceinline__ void fmm_pp_share(const double2* __restrict__ pd, uint32_t q,
                                             uint32_t end, uint32_t F, double px, double py,
                                             double& ax, double& ay, double& t) {
    for (; q + 3 * F < end; q += 4 * F) {
        const double2 a = pd[q], b = pd[q + F], c = pd[q + 2 * F], d = pd[q + 3 * F];
        fmm_pp1(a, px, py, ax, ay, t, true);
        fmm_pp1(b, px, py, ax, ay, t, true);
        fmm_pp1(c, px, py, ax, ay, t, true);
        fmm_pp1(d, px, py, ax, ay, t, true);
    }
    for (; q < end; q += F) fmm_pp1(pd[q], px, py, ax, ay, t, true);
}

// One unit of the leaf phase: chunk c0 of leaf X, share f of F. GROUPED (leaves of at most 32
// points): G groups of Wc lanes hold the same sinks and take G different source shares
// (share f G + g of F G), summed over the groups at the end.
template <bool GROUPED>
__device__ __forceinline__ void fmm_pp_unit(const double2* __restrict__ pd,
                                            const uint32_t* __restrict__ first,
                                            const uint32_t* __restrict__ count, int depth,
                                            FmmMoments M, double side, double theta, double thr,
                                            const uint32_t* __restrict__ lists,
                                            const uint32_t* __restrict__ nlist, uint32_t X,
                                            uint32_t ix, uint32_t F, uint32_t c0, uint32_t f,
                                            double2* __restrict__ stage, uint64_t part_stride,
                                            double* __restrict__ part, double& t_out,
                                            bool& mine_out) {
    const int lane = threadIdx.x & 63;
    const uint32_t fX = first[X], nX = count[X];
    const uint32_t Wc = GROUPED ? (nX > 16u ? 32u : (nX > 8u ? 16u : 8u)) : (uint32_t)kWave;
    const uint32_t G = GROUPED ? (uint32_t)kWave / Wc : 1u;
    const uint32_t g = GROUPED ? (uint32_t)lane / Wc : 0u;
    const uint32_t sl = GROUPED ? ((uint32_t)lane & (Wc - 1)) : (uint32_t)lane;
    const uint32_t FG = F * G, share = f * G + g;
    const bool mine = c0 + sl < nX;
    const uint32_t p = fX + c0 + (mine ? sl : 0);
    const double2 pp = pd[p];
    const double px = pp.x, py = pp.y;
    const double cxX = M.cx[X], cyX = M.cy[X];
    double ax = 0.0, ay = 0.0, t = 0.0;
    fmm_partners_meta(first, count, M, side, theta, depth, X, ix, lists, nlist,
                      [&](uint32_t Y, uint32_t fY, uint32_t end, double cxY, double cyY) {
        if (Y == X) {
            fmm_pp_share(pd, fY + share, end, FG, px, py, ax, ay, t);
            // the sink itself is in its own share once (w = 1, no force): take it out
            if ((p - fX) % FG == share) t -= 1.0;
        } else if (X < Y) {
            const double rx = px - cxY, ry = py - cyY;
            if (rx * rx + ry * ry > thr * thr) {
                if (share == 0) {
                    double rS, D1, D2, D3;
                    fmm_derivs(rx, ry, rS, D1, D2, D3);
                    double vx, vy;
                    fmm_c1(rx, ry, D1, D2, D3, M.m0[Y], M.xx[Y], M.xy[Y], M.yy[Y], vx, vy);
                    ax += vx;
                    ay += vy;
                    t += 2.0 * M.m0[Y] / rS;
                }
            } else {
                fmm_pp_share(pd, fY + share, end, FG, px, py, ax, ay, t);
            }
        } else if (!GROUPED) {
            // Y's points tested against X's box 64 at a time (one per lane), the ones kept
            // compacted into LDS and read back by the whole wave
            for (uint32_t b0 = fY + share; b0 < end; b0 += FG * (uint32_t)kWave) {
                const uint32_t q = b0 + FG * (uint32_t)lane;
                bool keep = false;
                double2 qq = make_double2(0.0, 0.0);
                if (q < end) {
                    qq = pd[q];
                    const double bx = qq.x - cxX, by = qq.y - cyX;
                    keep = !(bx * bx + by * by > thr * thr);
                }
                const uint64_t kb = __ballot(keep);
                if (keep) stage[__popcll(kb & lane_mask_lt())] = qq;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const uint32_t nk = (uint32_t)__popcll(kb);
                uint32_t j = 0;
                for (; j + 3 < nk; j += 4) {
                    const double2 a = stage[j], b = stage[j + 1], c = stage[j + 2],
                                  d = stage[j + 3];
                    fmm_pp1(a, px, py, ax, ay, t, true);
                    fmm_pp1(b, px, py, ax, ay, t, true);
                    fmm_pp1(c, px, py, ax, ay, t, true);
                    fmm_pp1(d, px, py, ax, ay, t, true);
                }
                for (; j < nk; ++j) fmm_pp1(stage[j], px, py, ax, ay, t, true);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
        } else {
            for (uint32_t q = fY + share; q < end; q += FG) {
                const double2 qq = pd[q];
                const double bx = qq.x - cxX, by = qq.y - cyX;
                if (bx * bx + by * by > thr * thr) continue;
                fmm_pp1(qq, px, py, ax, ay, t, true);
            }
        }
    });
    if (GROUPED) {
        for (uint32_t o = Wc; o < (uint32_t)kWave; o <<= 1) {
            ax += __shfl_xor(ax, (int)o, kWave);
            ay += __shfl_xor(ay, (int)o, kWave);
        }
    }
    if (mine && g == 0) {
        part[(size_t)(2 * f) * part_stride + p] = ax;
        part[(size_t)(2 * f + 1) * part_stride + p] = ay;
    }
    t_out = t;
    mine_out = mine;
}

// The same unit with its sources staged in LDS first (units of at most 64 partners whose
// shares hold at most kFmmStage points): lane k copies the share of partner k into its segment
// of the wave's stage buffer, dropping the points of a lower-slot partner that X's box accepts,
// so every load of the unit is in flight at once; the sums then read the sources from LDS.
// Returns false (nothing done) when the unit does not qualify.
template <bool GROUPED>
__device__ __forceinline__ bool fmm_pp_unit_staged(const double2* __restrict__ pd,
                                                   const uint32_t* __restrict__ first,
                                                   const uint32_t* __restrict__ count,
                                                   FmmMoments M, double thr,
                                                   const uint32_t* __restrict__ lists,
                                                   const uint32_t* __restrict__ nlist, uint32_t X,
                                                   uint32_t F, uint32_t c0, uint32_t f,
                                                   double2* __restrict__ stage,
                                                   uint64_t part_stride, double* __restrict__ part,
                                                   double& t_out, bool& mine_out) {
    const int lane = threadIdx.x & 63;
    const uint32_t np = nlist[X];
    if (np == kFmmOverflow || np > (uint32_t)kWave) return false;
    const uint32_t fX = first[X], nX = count[X];
    const double cxX = M.cx[X], cyX = M.cy[X];
    // partner k in lane k
    const bool has = (uint32_t)lane < np;
    const uint32_t Y = has ? lists[(size_t)X * kFmmListCap + lane] : X;
    const uint32_t fY = first[Y], nY = has ? count[Y] : 0u;
    const uint32_t m = nY > f ? (nY - 1 - f) / F + 1 : 0u;  // points of share f
    uint32_t off = m;  // exclusive prefix over the lanes
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t y = __shfl_up(off, d, kWave);
        if (lane >= d) off += y;
    }
    const uint32_t all = (uint32_t)__builtin_amdgcn_readlane((int)off, kWave - 1);
    if (all > kFmmStage) return false;
    off -= m;
    const bool second = Y < X;
    uint32_t kept = 0;
    for (uint32_t i = 0; i < m; i += 4) {
        double2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = pd[fY + f + F * (i + u < m ? i + u : i)];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (i + u >= m) break;
            if (second) {
                const double bx = v[u].x - cxX, by = v[u].y - cyX;
                if (bx * bx + by * by > thr * thr) continue;
            }
            stage[off + kept] = v[u];
            ++kept;
        }
    }
    const double cxl = M.cx[Y], cyl = M.cy[Y];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    const uint32_t Wc = GROUPED ? (nX > 16u ? 32u : (nX > 8u ? 16u : 8u)) : (uint32_t)kWave;
    const uint32_t G = GROUPED ? (uint32_t)kWave / Wc : 1u;
    const uint32_t g = GROUPED ? (uint32_t)lane / Wc : 0u;
    const uint32_t sl = GROUPED ? ((uint32_t)lane & (Wc - 1)) : (uint32_t)lane;
    const bool mine = c0 + sl < nX;
    const uint32_t o = c0 + (mine ? sl : 0);  // sink offset in the leaf
    const uint32_t p = fX + o;
    const double2 pp = pd[p];
    const double px = pp.x, py = pp.y;
    double ax = 0.0, ay = 0.0, t = 0.0;
    for (uint32_t k = 0; k < np; ++k) {
        const uint32_t Yk = (uint32_t)__builtin_amdgcn_readlane((int)Y, (int)k);
        const uint32_t sb = (uint32_t)__builtin_amdgcn_readlane((int)off, (int)k);
        const uint32_t se = sb + (uint32_t)__builtin_amdgcn_readlane((int)kept, (int)k);
        bool run = true;
        if (X < Yk) {
            const double rx = px - fmm_readlane_d(cxl, k), ry = py - fmm_readlane_d(cyl, k);
            if (rx * rx + ry * ry > thr * thr) {
                run = false;
                if (f == 0 && g == 0) {
                    double rS, D1, D2, D3;
                    fmm_derivs(rx, ry, rS, D1, D2, D3);
                    double vx, vy;
                    fmm_c1(rx, ry, D1, D2, D3, M.m0[Yk], M.xx[Yk], M.xy[Yk], M.yy[Yk], vx, vy);
                    ax += vx;
                    ay += vy;
                    t += 2.0 * M.m0[Yk] / rS;
                }
            }
        }
        if (!run) continue;
        uint32_t j = sb + g;
        for (; j + 3 * G < se; j += 4 * G) {
            const double2 a = stage[j], b = stage[j + G], c = stage[j + 2 * G], d = stage[j + 3 * G];
            fmm_pp1(a, px, py, ax, ay, t, true);
            fmm_pp1(b, px, py, ax, ay, t, true);
            fmm_pp1(c, px, py, ax, ay, t, true);
            fmm_pp1(d, px, py, ax, ay, t, true);
        }
        for (; j < se; j += G) fmm_pp1(stage[j], px, py, ax, ay, t, true);
        // the leaf's own share holds the sink itself once (w = 1, no force): take it out
        if (Yk == X && o % F == f && (o / F) % G == g) t -= 1.0;
    }
    if (GROUPED) {
        for (uint32_t w = Wc; w < (uint32_t)kWave; w <<= 1) {
            ax += __shfl_xor(ax, (int)w, kWave);
            ay += __shfl_xor(ay, (int)w, kWave);
        }
    }
    if (mine && g == 0) {
        part[(size_t)(2 * f) * part_stride + p] = ax;
        part[(size_t)(2 * f + 1) * part_stride + p] = ay;
    }
    t_out = t;
    mine_out = mine;
    // the next unit's staging must not overwrite sources still being read
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    return true;
}

// The leaf phase of the pair recursion, one unit per wave at a time: its sinks of leaf X against
// every leaf partner Y (X itself included). The reference orients a leaf pair (A, B) with A
// the lower slot; only A's points are tested against B's box (l_B/theta):
//  - Y == X: point-point over the leaf's other points;
//  - X < Y (X first): a sink far enough from Y's centre takes Y's expansion (C1 of the
//    point-node kernel, share 0 only), the others sum over Y's points;
//  - Y < X (X second): Y's points not accepted against X's box act point by point on the sinks
//    (the accepted ones went into X's expansion in k_fmm_exp).
// Unit f takes partner points f, f + F, ... (split further over lane groups for small leaves);
// the F shares are summed per point afterwards.
template <bool LIGHT>
__global__ __launch_bounds__(kBlock) void k_fmm_pp(const double2* __restrict__ pd,
                                                   const uint32_t* __restrict__ first,
                                                   const uint32_t* __restrict__ count, int depth,
                                                   uint32_t S, const BuildParams* __restrict__ prm,
                                                   double theta, double* __restrict__ mom,
                                                   const uint32_t* __restrict__ lists,
                                                   const uint32_t* __restrict__ nlist,
                                                   const uint32_t* __restrict__ unit_off,
                                                   const uint32_t* __restrict__ split,
                                                   uint64_t part_stride,
                                                   double* __restrict__ part,
                                                   double* __restrict__ tot_unit) {
    const int lane = threadIdx.x & 63;
    const uint32_t L = 1u << (2 * depth);
    const uint32_t total = unit_off[L];
    const uint32_t waves = gridDim.x * kWavesPerBlock;
    FmmMoments M = fmm_moments(mom, S);
    double side = __longlong_as_double((long long)prm->side64);
    if (!(side > 0.0)) side = 1.0;
    const double thr = ldexp(side, -depth) / theta;
    const uint32_t base = fmm_level_index(depth);
    __shared__ double2 stage_all[kWavesPerBlock][LIGHT ? kFmmStage : kWave];
    for (uint32_t u = __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
         u < total; u += waves) {
        uint32_t lo = 0, hi = L;  // unit_off[lo] <= u < unit_off[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (unit_off[mid] <= u) lo = mid;
            else hi = mid;
        }
        const uint32_t ix = lo, X = base + ix;
        const uint32_t F = split[ix], local = u - unit_off[ix];
        const uint32_t c0 = (local / F) * kWave, f = local % F;
        double t;
        bool mine;
        const bool grouped = count[X] <= 32u;
        bool done = false;
        double2* stage = stage_all[threadIdx.x >> 6];
        if (LIGHT) {
            done = grouped ? fmm_pp_unit_staged<true>(pd, first, count, M, thr, lists, nlist, X, F,
                                                      c0, f, stage, part_stride, part, t, mine)
                           : fmm_pp_unit_staged<false>(pd, first, count, M, thr, lists, nlist, X, F,
                                                       c0, f, stage, part_stride, part, t, mine);
        }
        if (!done) {
            if (grouped)
                fmm_pp_unit<true>(pd, first, count, depth, M, side, theta, thr, lists, nlist, X, ix,
                                  F, c0, f, stage, part_stride, part, t, mine);
            else
                fmm_pp_unit<false>(pd, first, count, depth, M, side, theta, thr, lists, nlist, X, ix,
                                   F, c0, f, stage, part_stride, part, t, mine);
        }
        t = wave_sum_d(mine ? t : 0.0);
        if (lane == 0) tot_unit[u] = t;
    }
}

// applyForces (:482-536), one level: C2, C3 of a node divided by its M0, then the parent's
// (final) expansion translated to the node's centre of mass added. Level 0: the division only.
__device__ __forceinline__ void fmm_down_node(const uint32_t* __restrict__ count, int l,
                                              uint32_t i, uint32_t S, double* __restrict__ mom,
                                              double* __restrict__ C) {
    const uint32_t s = fmm_level_index(l) + i;
    if (count[s] == 0u) return;
    FmmMoments M = fmm_moments(mom, S);
    FmmCoef c = fmm_load(C, S, s);
    const double inv = 1.0 / M.m0[s];
#pragma unroll
    for (int k = 0; k < 4; ++k) c.c2[k] *= inv;
#pragma unroll
    for (int k = 0; k < 8; ++k) c.c3[k] *= inv;
    if (l > 0) {
        const uint32_t par = fmm_level_index(l - 1) + (i >> 2);
        const FmmCoef pc = fmm_load(C, S, par);
        double vx, vy, t2[4];
        fmm_translate(pc, M.cx[s] - M.cx[par], M.cy[s] - M.cy[par], vx, vy, t2);
        c.c1[0] += vx;
        c.c1[1] += vy;
#pragma unroll
        for (int k = 0; k < 4; ++k) c.c2[k] += t2[k];
#pragma unroll
        for (int k = 0; k < 8; ++k) c.c3[k] += pc.c3[k];
    }
    fmm_store(C, S, s, c);
}

__global__ __launch_bounds__(kBlock) void k_fmm_down(const uint32_t* __restrict__ count, int l,
                                                     uint32_t S, double* __restrict__ mom,
                                                     double* __restrict__ C) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < (1u << (2 * l))) fmm_down_node(count, l, i, S, mom, C);
}

__global__ __launch_bounds__(kFmmTopNodes) void k_fmm_down_top(const uint32_t* __restrict__ count,
                                                               int top, uint32_t S,
                                                               double* __restrict__ mom,
                                                               double* __restrict__ C) {
    for (int l = 0; l <= top; ++l) {
        if (threadIdx.x < (1u << (2 * l))) fmm_down_node(count, l, threadIdx.x, S, mom, C);
        __syncthreads();
    }
}

// evaluation at the points: derivative = the F_X point-point shares + local expansion of the
// leaf (FP32 out). One thread per point; its leaf is the top 2*depth bits of its Morton code.
__global__ __launch_bounds__(kBlock) void k_fmm_l2p(const float4* __restrict__ pos4,
                                                    const uint32_t* __restrict__ codes, uint32_t n,
                                                    int depth, uint32_t S,
                                                    const double* __restrict__ mom,
                                                    const double* __restrict__ C,
                                                    const uint32_t* __restrict__ split,
                                                    uint64_t part_stride,
                                                    const double* __restrict__ part,
                                                    float* __restrict__ ax, float* __restrict__ ay) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const uint32_t ix = depth == 0 ? 0u : codes[p] >> (2 * (16 - depth));
    const uint32_t X = fmm_level_index(depth) + ix;
    const uint32_t F = split[ix];
    double dx = 0.0, dy = 0.0;
    for (uint32_t f = 0; f < F; ++f) {
        dx += part[(size_t)(2 * f) * part_stride + p];
        dy += part[(size_t)(2 * f + 1) * part_stride + p];
    }
    const FmmCoef c = fmm_load(C, S, X);
    const float4 pp = pos4[p];
    double vx, vy;
    fmm_translate(c, (double)pp.x - mom[(size_t)S + X], (double)pp.y - mom[2 * (size_t)S + X], vx,
                  vy, nullptr);
    ax[p] = (float)(dx + vx);
    ay[p] = (float)(dy + vy);
}

// Z: sum of the node-node totals (S slots) and the light and heavy leaf-unit totals in a fixed
// order (block partials, then one block)
__global__ __launch_bounds__(kBlock) void k_fmm_sum_partial(const double* __restrict__ a,
                                                            uint32_t na,
                                                            const double* __restrict__ b,
                                                            const uint32_t* __restrict__ nb_ptr,
                                                            const double* __restrict__ c,
                                                            const uint32_t* __restrict__ nc_ptr,
                                                            double* __restrict__ partials) {
    const uint32_t nb = *nb_ptr, nc = *nc_ptr;
    const uint32_t step = gridDim.x * kBlock;
    double s = 0.0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < na; i += step) s += a[i];
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < nb; i += step) s += b[i];
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < nc; i += step) s += c[i];
    s = wave_sum_d(s);
    __shared__ double red[kWavesPerBlock];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < kWavesPerBlock; ++k) t += red[k];
        partials[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(64) void k_fmm_sum_final(const double* __restrict__ partials,
                                                      uint32_t nb, double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double t = 0.0;
    for (uint32_t k = 0; k < nb; ++k) t += partials[k];
    *out = t;
}

static inline uint32_t fmm_div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

uint64_t fmm_slots(int depth) { return (((uint64_t)1 << (2 * (depth + 1))) - 1) / 3; }

namespace {
constexpr uint32_t kFmmSumBlocks = 64;

// workspace layout (each part 256-byte aligned): moments 6 x S, coefficients 14 x S, node
// totals S, unit totals 2 x max_units (light, heavy), point-point shares 2 x kFmmMaxSplit x n_max,
// FP64 positions n_max, Z partials, unit offsets 3 x (S + 1 rounded up to 4) (light, heavy,
// expansion), splits S, list lengths S, lists kFmmListCap x S (uint32). The expansion partials (14 per expansion
// unit, at most 14 (leaves + n/4) doubles) live in the point-point shares' space, which is
// written only after they are summed.
struct FmmLayout {
    uint64_t mom, coef, tot_nn, tot_unit, part, pd, partials, units, split, nlist, lists, bytes,
        max_units, unit_stride;
};

FmmLayout fmm_layout(uint64_t S, uint64_t n_max) {
    auto up = [](uint64_t b) { return (b + 255) & ~(uint64_t)255; };
    FmmLayout L{};
    const uint64_t leaves = S < n_max ? S : n_max;
    L.max_units = (uint64_t)kFmmMaxSplit * ((n_max + kWave - 1) / kWave + leaves);
    uint64_t o = 0;
    L.mom = o;       o += up(6 * S * sizeof(double));
    L.coef = o;      o += up(14 * S * sizeof(double));
    L.tot_nn = o;    o += up(S * sizeof(double));
    L.tot_unit = o;  o += up(2 * L.max_units * sizeof(double));
    L.part = o;      o += up(2ull * kFmmMaxSplit * n_max * sizeof(double));
    L.pd = o;        o += up(n_max * sizeof(double2));
    L.partials = o;  o += up(kFmmSumBlocks * sizeof(double));
    L.unit_stride = (S + 1 + 3) & ~(uint64_t)3;
    L.units = o;     o += up(3 * L.unit_stride * sizeof(uint32_t));
    L.split = o;     o += up(S * sizeof(uint32_t));
    L.nlist = o;     o += up(S * sizeof(uint32_t));
    L.lists = o;     o += up((uint64_t)kFmmListCap * S * sizeof(uint32_t));
    L.bytes = o;
    return L;
}
}  // namespace

uint64_t fmm_workspace_bytes(uint64_t slot_stride, uint64_t n_max) {
    return fmm_layout(slot_stride, n_max).bytes;
}

void launch_fmm(hipStream_t s, const float4* pos4, const uint32_t* codes, const uint32_t* first,
                const uint32_t* count, int depth, uint64_t slot_stride, uint32_t n, uint64_t n_max,
                const BuildParams* prm, double theta, void* workspace, float* ax, float* ay,
                double* total) {
    const uint32_t S = (uint32_t)slot_stride;
    const FmmLayout Lw = fmm_layout(slot_stride, n_max);
    char* w = (char*)workspace;
    double* mom = (double*)(w + Lw.mom);
    double* coef = (double*)(w + Lw.coef);
    double* tot_nn = (double*)(w + Lw.tot_nn);
    double* tot_unit = (double*)(w + Lw.tot_unit);
    double* part = (double*)(w + Lw.part);
    double* partials = (double*)(w + Lw.partials);
    uint32_t* units = (uint32_t*)(w + Lw.units);
    uint32_t* split = (uint32_t*)(w + Lw.split);
    uint32_t* nlist = (uint32_t*)(w + Lw.nlist);
    const uint32_t US = (uint32_t)Lw.unit_stride;
    uint32_t* lists = (uint32_t*)(w + Lw.lists);
    const uint32_t leaves = 1u << (2 * depth);
    const uint32_t leaf_waves = fmm_div_up(leaves, kWavesPerBlock);
    double2* pd = (double2*)(w + Lw.pd);
    k_fmm_leaf_moments<<<leaf_waves, kBlock, 0, s>>>(pos4, first, count, depth, S, mom, pd);
    const int top = depth - 1 < kFmmTopLevel ? depth - 1 : kFmmTopLevel;
    for (int l = depth - 1; l > top; --l)
        k_fmm_up<<<fmm_div_up(1u << (2 * l), kBlock), kBlock, 0, s>>>(l, S, mom);
    if (top >= 0) k_fmm_up_top<<<1, kFmmTopNodes, 0, s>>>(top, S, mom);
    for (int l = 0; l <= depth; ++l)
        k_fmm_level<<<fmm_div_up(1u << (2 * l), kWavesPerBlock), kBlock, 0, s>>>(
            count, l, depth, S, prm, theta, mom, coef, tot_nn, lists, nlist, units, US, split);
    k_fmm_scan<<<1, kFmmScanThreads, 0, s>>>(units, US, leaves);
    {
        const uint64_t ebound = (uint64_t)(leaves < n ? leaves : n) + n / 4;
        uint32_t eblocks = fmm_div_up(ebound, kWavesPerBlock);
        if (eblocks > kFmmUnitBlocks) eblocks = kFmmUnitBlocks;
        if (eblocks < 1) eblocks = 1;
        k_fmm_exp<<<eblocks, kBlock, 0, s>>>(pd, first, count, depth, S, prm, theta, mom, lists,
                                             nlist, units + 2 * US, part);
        k_fmm_exp_add<<<fmm_div_up(leaves, kBlock), kBlock, 0, s>>>(units + 2 * US, depth, S,
                                                                    part, coef);
    }
    // the unit count stays on the device: a fixed grid of waves strides over the units
    const uint64_t unit_bound = (uint64_t)kFmmMaxSplit * ((n + kWave - 1) / kWave +
                                                          (leaves < n ? leaves : n));
    uint32_t blocks = fmm_div_up(unit_bound, kWavesPerBlock);
    if (blocks > kFmmUnitBlocks) blocks = kFmmUnitBlocks;
    if (blocks < 1) blocks = 1;
    k_fmm_pp<true><<<blocks, kBlock, 0, s>>>(pd, first, count, depth, S, prm, theta, mom, lists,
                                             nlist, units, split, n_max, part, tot_unit);
    k_fmm_pp<false><<<blocks, kBlock, 0, s>>>(pd, first, count, depth, S, prm, theta, mom, lists,
                                              nlist, units + US, split, n_max, part,
                                              tot_unit + Lw.max_units);
    const int dtop = depth < kFmmTopLevel ? depth : kFmmTopLevel;
    k_fmm_down_top<<<1, kFmmTopNodes, 0, s>>>(count, dtop, S, mom, coef);
    for (int l = dtop + 1; l <= depth; ++l)
        k_fmm_down<<<fmm_div_up(1u << (2 * l), kBlock), kBlock, 0, s>>>(count, l, S, mom, coef);
    k_fmm_l2p<<<fmm_div_up(n, kBlock), kBlock, 0, s>>>(pos4, codes, n, depth, S, mom, coef, split,
                                                       n_max, part, ax, ay);
    const uint32_t used = (uint32_t)fmm_slots(depth);
    k_fmm_sum_partial<<<kFmmSumBlocks, kBlock, 0, s>>>(tot_nn, used, tot_unit, units + leaves,
                                                       tot_unit + Lw.max_units,
                                                       units + US + leaves, partials);
    k_fmm_sum_final<<<1, 64, 0, s>>>(partials, kFmmSumBlocks, total);
}

}  // namespace bh
