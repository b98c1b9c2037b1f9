// Host-side launchers of the Barnes-Hut kernels (defined next to their kernels).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bh_common.h"

namespace bh {

constexpr int kMaxLeafCap = 256;       // largest leaf capacity of the adaptive build
constexpr uint32_t kMaxBoundsBlocks = 1024;
constexpr uint32_t kMaxReduceBlocks = 1024;
constexpr uint32_t kStatsBlocks = 256;
constexpr int kKernelGravity = 0;
constexpr int kKernelTsne = 1;

uint32_t bounds_grid(uint32_t n);
// Build status words (device, cleared by the host before a build sequence, read after it):
// [0] the radix sort's look-back gave up waiting, [1] a position was not finite.
constexpr int kStatusWords = 4;
constexpr uint32_t kSortSpinLimit = 1u << 22;  // look-back polls before giving up
void launch_bounds(hipStream_t s, const float* xs, const float* ys, uint32_t n, float4* partials,
                   BuildParams* prm, uint32_t tree_mode, const double* caller, uint32_t* status);
// keys of the reference GPU step's tree (32-bit, path in the top 2 depth bits)
void launch_morton_gpuref(hipStream_t s, const float* xs, const float* ys, uint32_t n,
                          const BuildParams* prm, int depth, uint32_t* keys, uint32_t* vals);
void launch_morton(hipStream_t s, const float* xs, const float* ys, uint32_t n,
                   const BuildParams* prm, int B, void* keys, uint32_t* vals, bool wide);
uint32_t scan_partials_size(uint64_t n);
void launch_exclusive_scan(hipStream_t s, const uint32_t* in, uint32_t* out, uint32_t n,
                           uint32_t* partials, uint32_t* total);
uint32_t sort_blocks(uint32_t n);
uint64_t sort_scratch_words(uint32_t n);  // uint32 words of scratch launch_radix_sort needs
// err: status word set when a look-back wait exceeds spin_limit polls (the order is then not
// valid; later passes do nothing and launch_gather falls back to the input order)
void launch_radix_sort(hipStream_t s, void* keys_a, uint32_t* vals_a, void* keys_b,
                       uint32_t* vals_b, uint32_t n, int key_bits, bool wide, uint32_t* scratch,
                       uint32_t* err, uint32_t spin_limit, bool* result_in_b);
// codes: the sorted codes, zeroed (with the identity order) when *sort_err is set
void launch_gather(hipStream_t s, const uint32_t* perm, uint32_t n, const float* x, const float* y,
                   const float* m, const float* vx, const float* vy, const uint32_t* id, float* xo,
                   float* yo, float* mo, float* vxo, float* vyo, uint32_t* ido, float4* pos4,
                   const uint32_t* sort_err, void* codes, bool wide);
// R: level of the tree roots (0; batched systems: the system id occupies the levels above R)
void launch_tree(hipStream_t s, const void* codes, bool wide, uint32_t n, int B, int cap,
                 int fixed_depth, int single_child_pp, uint8_t* leaf_lvl, uint8_t* lcp1,
                 uint32_t* cnt, uint32_t* noff, uint32_t* scan_partials, BuildParams* prm,
                 Node* nodes, uint32_t* parent, uint32_t* child4, uint32_t cap_nodes, int R,
                 const BuildParams* sys_prm, const uint32_t* status);
void launch_bounds_batched(hipStream_t s, const float* xs, const float* ys, const uint2* blk,
                           uint32_t nblk, const uint32_t* blk_off, const uint32_t* sys_off,
                           uint32_t nsys, float4* partials, BuildParams* sys_prm,
                           BuildParams* prm, uint32_t* status);
void launch_morton_batched(hipStream_t s, const float* xs, const float* ys, uint32_t n,
                           const uint32_t* sys_off, uint32_t nsys, const BuildParams* sys_prm,
                           int B, uint64_t* keys, uint32_t* vals);
void launch_batch_roots(hipStream_t s, const uint32_t* sys_off, uint32_t nsys,
                        const uint32_t* noff, const BuildParams* prm, uint32_t* roots);
uint32_t up_chunks(uint32_t cap_nodes);
// deferred and bylevel: up_chunks(cap_nodes) * kDeferStride words each
void launch_upward(hipStream_t s, const float4* pos4, Node* nodes, const uint32_t* child4,
                   float4* m2, const BuildParams* prm, uint32_t* deferred, uint32_t* bylevel,
                   uint32_t cap_nodes, bool big_leaves);
void launch_tree_stats(hipStream_t s, const Node* nodes, const BuildParams* prm,
                       uint32_t cap_nodes, int depth_limit, uint32_t leaf_cap, uint2* out,
                       uint32_t blocks);
void launch_fill_u32(hipStream_t s, uint32_t* p, uint64_t n, uint32_t v);
void launch_fill_f32(hipStream_t s, float* p, uint64_t n, float v);
void launch_iota(hipStream_t s, uint32_t* p, uint32_t n);
void launch_tree_slots(hipStream_t s, const uint32_t* codes, const Node* nodes,
                       const BuildParams* prm, uint32_t cap_nodes, int depth, uint32_t* first_child,
                       uint32_t* first_particle, uint32_t* count, uint64_t slots);

// Traversal variants: sinks per lane NS in {1, 2, 4}, monopoles evaluated only when some lane
// accepts (branchy) or always under a mask, and the particle-particle source fetch: PPU sources
// per scalar load (1 or 4), or PPU = 0: each lane loads one of up to 64 sources with one vector
// load and the wave broadcasts them with readlane (one memory round trip per cell).
// Variants 0..11: NS = {1,2,4}[v / 4], branchy = (v % 4) < 2, PPU = v % 2 ? 4 : 1;
// variants 12..17: NS = {1,2,4}[(v - 12) / 2], branchy = (v - 12) % 2 == 0, PPU = 0;
// variants 18..35: the shapes of v - 18 with the lane masks in scalar registers.
constexpr int kTraverseVariantsVector = 18;
constexpr int kTraverseVariants = 2 * kTraverseVariantsVector;
inline int variant_shape(int v) { return v % kTraverseVariantsVector; }
inline int variant_ns_index(int v) {
    v = variant_shape(v);
    return v < 12 ? v / 4 : (v - 12) / 2;
}
inline int variant_ns(int v) { return 1 << variant_ns_index(v); }
inline bool variant_branchy(int v) {
    v = variant_shape(v);
    return v < 12 ? (v % 4) < 2 : (v - 12) % 2 == 0;
}
inline int variant_ppu(int v) {
    v = variant_shape(v);
    return v < 12 ? (v % 2 ? 4 : 1) : 0;
}
// returns the number of blocks launched (one counter partial each); prio: the waves of a
// one-system walk that fits on the device at once lower their issue priority as they progress
// (scheduling only, results are identical either way)
// batched systems: one work entry (first sink, count, system, -) per wave and the systems'
// roots; work == nullptr for a one-system context
// A batched work entry (first sink, count, system, w). w = 0: one wave walks the entry and
// writes its sinks' results. w != 0: the entry is one of F slices of a heavy entry (F > 1): bits
// 0-3 the slice f, 4-7 F, 8-31 the heavy slot; its wave takes the monopoles of every F-th visit
// and the f-th slice of every particle-particle cell, and writes partial results to
// side[slot][f][sink] (planes x, y, total), which k_sum_heavy adds in slice order.
constexpr uint32_t kHeavyMax = 8192;  // heavy entries of a work table
constexpr uint32_t kHeavySlices = 8;  // waves per heavy entry
constexpr uint32_t kHeavySinks = 256; // sinks per work entry, at most (64 lanes x 4 per lane)
struct TraverseBatch {
    const uint4* work = nullptr;
    uint32_t n_work = 0;
    const uint32_t* roots = nullptr;
    float* side = nullptr;    // 3 planes of side_slots x kHeavySlices x kHeavySinks
    uint32_t side_slots = 0;  // heavy slots `side` holds room for
};
// sums the slices of each heavy entry (heavy: (first sink, count, F, -) per slot) into the outputs
void launch_sum_heavy(hipStream_t s, const float* side, uint32_t side_slots, const uint4* heavy,
                      uint32_t nheavy, float* ax, float* ay, float* tot);
uint32_t launch_traverse(hipStream_t s, int kernel, int variant, const float4* pos4,
                     const Node* nodes, const BuildParams* prm, uint32_t n, float inv_theta2,
                     float eps2, float scale, float* ax, float* ay, float* tot,
                     unsigned long long* counters, TraverseBatch batch, uint32_t F = 1,
                     bool prio = true);
// F > 1 (one-system contexts): F waves per sink group, outputs are F parts of n each
constexpr uint32_t kMaxSplit = 8;
void launch_sum_parts(hipStream_t s, const float* part, uint32_t F, uint32_t n, float* out);
uint32_t reduce_grid(uint32_t n);
// out[0..3] = FP64 kinetic energy, potential energy (sum m phi / 2), momentum x, y; partials
// holds 4 doubles per block (at most kMaxReduceBlocks / 4 blocks)
void launch_state_sums(hipStream_t s, const float* m, const float* vx, const float* vy,
                       const float* phi, uint32_t n, double* partials, double* out);
void launch_reduce_f64(hipStream_t s, const float* v, const float* w, uint32_t n, double* partials,
                       double* out);
uint32_t traverse_blocks(uint32_t n);
// symmetric dual-tree FMM with the t-SNE kernels (FP64) over the fixed-depth slot tree: slot
// arrays have slot_stride entries (>= fmm_slots(depth)); codes are the sorted Morton codes;
// workspace holds fmm_workspace_bytes(slot_stride, n_max) bytes. Writes ax, ay (sorted order,
// FP32) and *total (Z).
uint64_t fmm_slots(int depth);
uint64_t fmm_workspace_bytes(uint64_t slot_stride, uint64_t n_max);
void launch_fmm(hipStream_t s, const float4* pos4, const uint32_t* codes, const uint32_t* first,
                const uint32_t* count, int depth, uint64_t slot_stride, uint32_t n, uint64_t n_max,
                const BuildParams* prm, double theta, void* workspace, float* ax, float* ay,
                double* total);
// streams the node records and sorted positions into every XCD's L2 when they fit (else no-op)
void launch_warm_l2(hipStream_t s, const Node* nodes, const BuildParams* prm, const float4* pos4,
                    uint32_t n, uint32_t* sink);
void launch_unsort2(hipStream_t s, const uint32_t* id, uint32_t n, const float* a, const float* b,
                    float* ao, float* bo);
void launch_kick_drift(hipStream_t s, float* x, float* y, float* vx, float* vy, const float* ax,
                       const float* ay, uint32_t n, float hdt, float dt);
void launch_kick(hipStream_t s, float* vx, float* vy, const float* ax, const float* ay, uint32_t n,
                 float hdt);

uint32_t direct_splits(uint32_t n, uint32_t ns);
// sys_off (nsys + 1 offsets, may be null): sources restricted to the sink's own system
// outputs (any may be null) indexed like sinks_caller; an index >= n gives NaN
void launch_direct_f64(hipStream_t s, int kernel, const float4* pos4, const uint32_t* id,
                       uint32_t n, const uint32_t* sinks_caller, uint32_t ns, double G, double eps2,
                       uint32_t* where, uint32_t* sinks_sorted, double* part, uint32_t S,
                       double* out_x, double* out_y, double* out_t,
                       const uint32_t* sys_off = nullptr, uint32_t nsys = 0);
// the same with the sorted sinks and partials in the library's static device array (no
// allocation): at most kDirectStaticSinks sinks, and direct_static_fits(n, ns) must hold
constexpr uint32_t kDirectStaticSinks = 65536;
constexpr int kMaxDevices = 64;
// device address of this GPU's copy of the static array (hipGetSymbolAddress once per device,
// with `device` current)
hipError_t direct_static_scratch(int device, void** ptr);
bool direct_static_fits(uint32_t n, uint32_t ns);
void launch_direct_f64_static(hipStream_t s, int kernel, const float4* pos4, const uint32_t* id,
                              uint32_t n, const uint32_t* sinks_caller, uint32_t ns, double G,
                              double eps2, uint32_t* where, void* scratch_ptr, double* out_x,
                              double* out_y, double* out_t, const uint32_t* sys_off,
                              uint32_t nsys);

// stage also writes the interleaved copy xy that the update's neighbour gathers read
void launch_tsne_stage(hipStream_t s, const float* yx, const float* yy, uint32_t n, float* x,
                       float* y, float* m, uint32_t* id, float2* xy);
void launch_tsne_attract(hipStream_t s, const uint32_t* row_ptr, const uint32_t* col,
                         const float* val, uint32_t n, const float2* xy, float exag, float* att_x,
                         float* att_y);
void launch_tsne_combine(hipStream_t s, uint32_t n, const float* yx, const float* yy,
                         const float* pyx, const float* pyy, const float* rep_x,
                         const float* rep_y, const float* att_x, const float* att_y,
                         const double* z, double rep_scale, float lr, float mu, float* nyx,
                         float* nyy, float* dx, float* dy);
uint32_t tsne_kl_blocks(uint32_t n);
void launch_tsne_kl(hipStream_t s, const uint32_t* row_ptr, const uint32_t* col, const float* val,
                    uint32_t n, const float* yx, const float* yy, const double* z,
                    double* partials, double* out);
void launch_sum_f64(hipStream_t s, const double* v, uint32_t nv, double* out);

// learned far-field correction (bh_mlp.hip)
struct MlpBf16;
void launch_mlp_prep(hipStream_t s, const float* params, MlpBf16* w);
void launch_mlp_infer_bf16(hipStream_t s, const float* X, uint32_t n, const MlpBf16* w, float* Y);
void launch_mlp_forward_f32(hipStream_t s, const float* X, uint32_t n, const float* p, float* Y);
size_t mlp_scratch_floats(uint32_t B);
uint32_t mlp_grad_chunks(uint32_t B);
uint32_t mlp_loss_blocks(uint32_t B);
// (terms: basis terms the loss uses, 3 or 7)
void launch_mlp_grad(hipStream_t s, const float* X, const float* T, const uint32_t* order,
                     uint32_t B, const float* p, float* S, double* loss_part, float* partials,
                     int terms);
void launch_mlp_adam(hipStream_t s, float* p, const float* partials, uint32_t nchunks, float* m1,
                     float* m2, float lr, int step, float* grad_out, int apply);
void launch_mlp_adam_grad(hipStream_t s, float* p, const float* g, float* m1, float* m2, float lr,
                          int step);
void launch_collect_pairs(hipStream_t s, const float4* pos4, const Node* nodes,
                          const BuildParams* prm, const uint32_t* sinks, uint32_t ns,
                          float inv_theta2, float inv_band2, uint32_t cap, uint32_t* pairs,
                          uint32_t* counts);
void launch_sinks_sorted(hipStream_t s, const uint32_t* id, uint32_t n, const uint32_t* sinks_caller,
                         uint32_t ns, uint32_t* where, uint32_t* sinks_sorted);
// the walk with the learned correction from the per-node coefficient records
// (terms: 3 or 7 basis terms of the correction, as the network was trained)
uint32_t launch_traverse_learned(hipStream_t s, int variant, const float4* pos4, const Node* nodes,
                                 const float4* coef, const BuildParams* prm, uint32_t n,
                                 float inv_theta2, float inv_band2, float eps2, float scale,
                                 int terms, float* ax, float* ay, float* tot,
                                 unsigned long long* counters);
void launch_label_pairs(hipStream_t s, const float4* pos4, const Node* nodes, const float* feat,
                        const BuildParams* prm, const uint32_t* sinks, const uint32_t* pairs, const uint32_t* counts,
                        const uint32_t* offsets, uint32_t ns, uint32_t cap, double eps2, float* X,
                        float* T, uint32_t total);
// cell features (kMlpIn per node, bh_mlp.h) of the last build and the network's coefficient
// records (kCoefStride floats per node) from them
// (lists: cell_features_lists_words(n) words of scratch, n = the particle capacity)
void launch_cell_features(hipStream_t s, const float4* pos4, const Node* nodes, const float4* m2,
                          const void* codes, bool wide, int B, const BuildParams* prm,
                          uint32_t cap_nodes, float* feat, uint32_t* lists, uint32_t n);
uint32_t cell_list_cap(uint32_t n, uint32_t c);
size_t cell_features_lists_words(uint32_t n);
void launch_cell_coefs(hipStream_t s, const float* feat, const Node* nodes,
                       const BuildParams* prm, uint32_t cap_nodes, const MlpBf16* w,
                       float4* coef);

// t-SNE input affinities (bh_affinity.hip): FP16 copy padded to aff_dim_padded(dim) halfs per
// point, candidate buffers of aff_cap() slots per point, exact k nearest by FP64 distance,
// bandwidths by bisection, symmetrisation keyed by row * n + column
uint32_t aff_dim_padded(uint32_t dim);
uint32_t aff_cap();
void launch_aff_pack(hipStream_t s, const float* X, uint32_t n, uint32_t dim, float scale,
                     _Float16* Xh, float* nh, float* ah, float* eb);
void launch_aff_candidates(hipStream_t s, const _Float16* Xh, const float* nh, const float* ah,
                           const float* eb, uint32_t n, uint32_t dim, uint32_t k, uint32_t* cj,
                           float* clo, float* chi, uint32_t* ccount, uint32_t* overflow);
void launch_aff_refine(hipStream_t s, const float* X, uint32_t n, uint32_t dim, uint32_t k,
                       const uint32_t* cj, const uint32_t* ccount, uint32_t* idx, double* d2);
void launch_aff_bandwidth(hipStream_t s, const double* d2, uint32_t n, uint32_t k, double target,
                          double tol, int max_iter, double* P, double* beta);
void launch_sym_entries(hipStream_t s, const uint32_t* idx, uint32_t n, uint32_t k,
                        uint64_t* keys, uint32_t* vals);
void launch_sym_heads(hipStream_t s, const uint64_t* keys, uint32_t m, uint32_t* head);
void launch_sym_reduce(hipStream_t s, const uint64_t* keys, const uint32_t* vals,
                       const uint32_t* pos, uint32_t m, uint32_t n, uint32_t nk, const double* P,
                       uint32_t* col, float* val, uint32_t* row_ptr, const uint32_t* nnz);

}  // namespace bh
