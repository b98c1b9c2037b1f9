/*
 * CPU oracle for the Barnes-Hut force step — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU algorithm (LucasFurer/MasterThesisLucas), used as
 * the reference for checking the HIP path. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only to check results or as the timed CPU baseline, never as the product path.
 * Each function names the reference file:line it follows. Arithmetic is FP64 like the reference.
 *
 * Pinning (what checks each function beyond retyping it):
 *   orc_morton32/64      closed-form bit interleave vs brute force, hand boundary cases, and
 *                        exhaustive per-axis checks in tests/test_oracle.py
 *   orc_stable_sort_u64  invariants (non-decreasing keys, permutation, ties by index)
 *   orc_get_depth        the depth table quoted for N = 65,536 / 70,000 / 1,048,576 / 4,194,304
 *   orc_fixed_tree       brute-force recount of every slot on small inputs; root M0 = N,
 *                        root COM = mean, leaf M2 = direct second moment
 *   orc_adaptive_tree    brute-force cell counts; topology equal to orc_quadtree_build's
 *   orc_quadtree_*       bit-exact against the compiled reference (oracle/_ref) when present
 *   orc_bh_traverse      theta=0 equals the direct sum; one far cell equals one monopole;
 *                        against orc_quadtree_forces and oracle/_ref
 *   orc_direct           closed forms (two-body, symmetric rings), Newton's third law
 *   orc_tsne_*           finite differences of the KL cost; Z invariant
 *   orc_mlp_*            finite differences; Adam closed form
 *   orc_fmm_sym          theta -> 0 equals the exact t-SNE sum; error vs the exact sum shrinks
 *                        with theta; two far clusters against the exact pair sum
 *   orc_gpuref_codes32   hand cases (centre and boundary points go low, the maximum goes high)
 *                        and a per-point descent by the same comparisons in tests/test_oracle.py
 *   orc_bh_traverse_f32  strict-threshold hand cases (FP32 expression and both FP64 rules),
 *                        theta = 0 equals the direct sum, equal to orc_bh_traverse when no
 *                        decision differs, a hand case where the FP64 rule disagrees
 *   orc_cell_features    one- and two-particle cells and a box offset in closed form
 *   orc_learned_basis,   the exact quadrupole coefficients reproduce the exact cell residual
 *   orc_pair_label       with error O(q^2) on a cluster with zero third moments and O(q)
 *                        otherwise; a one-particle cell has residual 0
 *   orc_learned_loss_grad finite differences
 *   orc_bh_traverse_learned zero coefficients equal orc_bh_traverse_f32; the exact quadrupole
 *                        coefficients cut the error against the direct sum
 */
#ifndef BH_ORACLE_H
#define BH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_KERNEL_GRAVITY = 0, ORC_KERNEL_TSNE = 1 };
/* acceptance / near-field semantics */
enum {
    ORC_SEM_QUADTREE = 0, /* nBodySolverBH.h:56-89: l/|d| < theta; children<=1 -> PP over all */
    ORC_SEM_MORTON = 1    /* nBodySolverFMM_SYM_MORTON.h:272-313: d^2 > (l/theta)^2; PP at leaves */
};

/* ---- the reference GPU step's tree: nBodySolverGpuBH.cu:78-200 ----
 * The cells' recursive partition (createRoot, createChildren) of the FP32 positions x, y to
 * `depth` levels; out: per point its path, two bits per level (x on the even bit: child
 * (x high) | (y high) << 1), in the top 2 depth bits of a 32-bit code. */
void orc_gpuref_codes32(const float* x, const float* y, int64_t n, int depth, uint32_t* code);

/* ---- Morton codes: nBodySolverFMM_SYM_MORTON.h:538-561 ---- */
void orc_morton32(const double* x, const double* y, int64_t n, double minx, double miny, double L,
                  uint32_t* code);
/* same rule with `bits` bits per axis (bits <= 32): q = min(floor(2^bits * RN((p-min)/L)), 2^bits-1) */
void orc_morton64(const double* x, const double* y, int64_t n, double minx, double miny, double L,
                  int bits, uint64_t* code);
void orc_quantize_axis(const double* v, int64_t n, double vmin, double L, int bits, uint64_t* q);

/* ---- sort: stable by (key, original index); perm[i] = original index of i-th element ---- */
void orc_stable_sort_u64(const uint64_t* key, int64_t n, uint32_t* perm);

/* ---- fixed-depth Morton tree: nBodySolverFMM_SYM_MORTON.h:106-151, 328-470 ---- */
int orc_get_depth(int max_children, int64_t n);
int64_t orc_fixed_tree_slots(int depth);
void orc_fixed_tree(const double* xs, const double* ys, int64_t n, int depth, double minx,
                    double miny, double maxx, double maxy, uint32_t* first_child,
                    uint32_t* first_particle, uint32_t* count, double* bb_cx, double* bb_cy,
                    double* bb_len, double* com_x, double* com_y, double* m0, double* m2);

/* ---- adaptive (leaf-capacity) tree over sorted Morton codes, pre-order layout ----
 * Cells at level k are the top 2k bits of a code with `bits` bits per axis. A cell is split
 * while it holds more than leaf_cap particles (quadtree.h:98) and level < bits.
 * Returns the node count, or -1 if cap_nodes is too small. */
int64_t orc_adaptive_tree(const uint64_t* codes_sorted, int64_t n, int bits, int leaf_cap,
                          int64_t cap_nodes, int32_t* level, uint32_t* first, uint32_t* count,
                          uint32_t* skip, uint32_t* nchild);
/* pre-order form of the fixed-depth tree: every non-empty cell of level <= depth */
int64_t orc_fixed_preorder(const uint32_t* codes_sorted, int64_t n, int depth, int64_t cap_nodes,
                           int32_t* level, uint32_t* first, uint32_t* count, uint32_t* skip,
                           uint32_t* nchild);
/* mass and centre of mass per pre-order node (plain FP64 sums over each node's range) */
void orc_node_moments(const double* xs, const double* ys, const double* ms, int64_t nnodes,
                      const uint32_t* first, const uint32_t* count, double* cx, double* cy,
                      double* mass, double* m2 /* 3 per node: xx, xy, yy (mass weighted) */);

/* ---- Barnes-Hut traversal over a pre-order tree (sorted particles) ----
 * nBodySolverBH.h:56-89 (semantics QUADTREE) / nBodySolverFMM_SYM_MORTON.h:272-313 (MORTON),
 * with the t-SNE kernels of nBodySolverBH.h:94-116 or the Plummer-softened gravity kernel.
 * Node side at level k = root_side * 2^-k. pp[k] marks nodes whose near field is PP.
 * sinks: sorted indices to evaluate (NULL = all n); outputs indexed like sinks. */
void orc_bh_traverse(const double* xs, const double* ys, const double* ms, int64_t n,
                     const int32_t* level, const uint32_t* first, const uint32_t* count,
                     const uint32_t* skip, const uint8_t* pp, const double* cx, const double* cy,
                     const double* mass, int64_t nnodes, double root_side, int semantics,
                     int kernel, double G, double eps2, double theta, const int64_t* sinks,
                     int64_t nsinks, double* out_ax, double* out_ay, double* out_tot,
                     uint64_t* counters /* [pn, pp] */, int nthreads);

/* ---- the same walk with single-precision acceptance decisions on given node records ----
 * Positions, masses and node records (cx, cy, mass) in single precision as the GPU holds them;
 * node side l = root_side * 2^-level, decision d2 > l2 * inv_theta2 with
 * l2 = RN32(l*l), d2 = fma32(dx, dx, RN32(dy*dy)), dx = RN32(cx - px): the expression of the
 * GPU traversal. Interactions are summed in FP64; abs_x/abs_y (may be NULL) receive per-sink
 * sums of |contribution|. counters = [pn, pp, mismatches]: when cx64/cy64 are given, the
 * decisions on the walk that the reference's FP64 rule of `semantics` (with side64, theta)
 * would take the other way. */
void orc_bh_traverse_f32(const float* xs, const float* ys, const float* ms, int64_t n,
                         const int32_t* level, const uint32_t* first, const uint32_t* count,
                         const uint32_t* skip, const uint8_t* pp, const float* cx, const float* cy,
                         const float* mass, int64_t nnodes, float root_side, float inv_theta2,
                         int kernel, double G, double eps2, const double* cx64, const double* cy64,
                         double side64, double theta, int semantics, const int64_t* sinks,
                         int64_t nsinks, double* out_ax, double* out_ay, double* out_tot,
                         double* abs_x, double* abs_y, uint64_t* counters, int nthreads);

/* ---- FP64 compensated direct sum (nBodySolverNaive.h:53-140), self excluded by index ---- */
void orc_direct(const double* x, const double* y, const double* m, int64_t n, int kernel, double G,
                double eps2, const int64_t* sinks, int64_t nsinks, double* out_ax, double* out_ay,
                double* out_tot, int nthreads);

/* ---- QuadTree restatement: quadtree.h:29-171 + nBodySolverBH.h:33-116 ----
 * Build from positions (input order), FP64 with the FP32 rounding of returned child
 * mass/COM (quadtree.h:96,146). Traversal accumulates into out_ax/out_ay (+=) and returns the
 * running total like solveNbody (single thread: bit-compatible order). */
typedef struct orc_quadtree orc_quadtree;
orc_quadtree* orc_quadtree_build(const double* x, const double* y, int64_t n, int max_children);
void orc_quadtree_free(orc_quadtree* t);
int64_t orc_quadtree_node_count(const orc_quadtree* t);
/* per node, pre-order (children in HH,HL,LH,LL order): depth, count, nchildren, low/high x/y,
 * total mass (float widened), com */
void orc_quadtree_export(const orc_quadtree* t, int32_t* depth, uint32_t* count, uint32_t* nchild,
                         double* lowx, double* lowy, double* highx, double* highy, double* mass,
                         double* comx, double* comy);
/* sorted occupant order of the leaves/pp cells is not needed by callers: occupant sets are
 * reported as min/max original index per node for topology comparisons */
void orc_quadtree_occupants(const orc_quadtree* t, uint32_t* node_first, uint32_t* occ_index);
double orc_quadtree_forces(const orc_quadtree* t, const double* x, const double* y, int kernel,
                           double G, double m, double eps2, double theta, double* out_ax,
                           double* out_ay, uint64_t* counters, int nthreads);

/* ---- integrators ---- */
/* kick-drift-kick leapfrog with BH forces (adaptive tree, QUADTREE semantics, or fixed depth
 * MORTON semantics), FP64 state, one tree rebuild per step (gravitysim.h:99-121 order). */
void orc_leapfrog(double* x, double* y, double* vx, double* vy, const double* m, int64_t n,
                  int steps, double dt, double G, double eps2, double theta, int semantics,
                  int leaf_cap, int bits, int nthreads);

/* ---- t-SNE: tsne.h:179-233, 287-301, 356-381, 383-460 ---- */
double orc_tsne_theta(double min_theta, double max_theta, int iteration);
/* attraction (adds to dx/dy): exag * sum_j P_ij w_ij (y_j - y_i) */
void orc_tsne_attraction(const double* x, const double* y, int64_t n, const uint32_t* row_ptr,
                         const uint32_t* col, const double* val, double exag, double* dx,
                         double* dy);
/* KL divergence sum_{i!=j} P_ij log2(P_ij/Q_ij) with exact FP64 Z (threaded) */
double orc_tsne_kl(const double* x, const double* y, int64_t n, const uint32_t* row_ptr,
                   const uint32_t* col, const double* val, int nthreads);
double orc_tsne_z(const double* x, const double* y, int64_t n, int nthreads);
/* one iteration: derivative (4/Z) BH repulsion (QuadTree, leaf capacity 16) + attraction, then
 * y_next = y + lr d + mu (y - y_prev); returns the Barnes-Hut Z */
double orc_tsne_step(const double* x, const double* y, const double* px, const double* py,
                     int64_t n, const uint32_t* row_ptr, const uint32_t* col, const double* val,
                     double theta, double exag, double lr, double mu, double* dx, double* dy,
                     double* nx, double* ny, int nthreads);

/* ---- MLP (10 -> H -> H -> 2, SiLU), FP64 ---- */
/* symmetric dual-tree FMM, t-SNE kernels, over orc_fixed_tree slots; returns Z */
double orc_fmm_sym(const double* xs, const double* ys, int64_t n, int depth,
                   const uint32_t* first_child, const uint32_t* first_particle,
                   const uint32_t* count, const double* bb_len, const double* com_x,
                   const double* com_y, const double* m0, const double* m2, double theta,
                   double* ax, double* ay);
void orc_mlp_forward(const double* X, int64_t B, int din, int h, int dout, const double* W1,
                     const double* b1, const double* W2, const double* b2, const double* W3,
                     const double* b3, double* Y);
/* mean-squared-error loss and parameter gradients */
double orc_mlp_loss_grad(const double* X, const double* T, int64_t B, int din, int h, int dout,
                         const double* W1, const double* b1, const double* W2, const double* b2,
                         const double* W3, const double* b3, double* gW1, double* gb1,
                         double* gW2, double* gb2, double* gW3, double* gb3);
void orc_adam(double* p, const double* g, double* mom1, double* mom2, int64_t n, double lr,
              double beta1, double beta2, double eps, int64_t t);

/* ---- learned far-field correction (no reference counterpart) ----
 * Cell features (ORC_CELL_FEATURES per node): mass-weighted second (3) and third (4) moments
 * about the centre of mass normalised by M l^2 and M l^3, the centre of mass's offset from the
 * cell's box centre / l (2), log2(count) / 16; the third moments are 0 for cells of more than
 * ORC_LEARN_T3_MAX particles. The network maps them to ORC_LEARNED_TERMS
 * complex coefficients c_j of the basis (w, conj w, w^3, q, q w^2, q conj(w)^2, q w^4) of the
 * pair offset (w = d/|d|, d = com - sink, q = l/|d|); the correction to a monopole is
 * (M l^2 / |d|^4) sum_j c_j b_j. */
#define ORC_CELL_FEATURES 10
#define ORC_LEARNED_TERMS 7
#define ORC_LEARN_T3_MAX 65536u
void orc_cell_features(const double* xs, const double* ys, const double* ms, int64_t nnodes,
                       const int32_t* level, const uint32_t* first, const uint32_t* count,
                       const uint64_t* codes_sorted, int bits, double minx, double miny,
                       double root_side, const double* cx, const double* cy, const double* mass,
                       double* f);
/* b: 14 doubles (re, im of the 7 basis functions) */
void orc_learned_basis(double ux, double uy, double q, double* b);
/* out[0..5]: label (exact cell sum - monopole) / scale (x, y), offset features w_x, w_y, q, and
 * scale = M l^2 / r^4; the sink `self` (sorted index, -1: none) is left out of the cell sum */
void orc_pair_label(const double* xs, const double* ys, const double* ms, uint32_t first,
                    uint32_t count, double cx, double cy, double mass, double l, double px,
                    double py, int64_t self, double eps2, double* out);
/* loss sum |sum_{j < terms} c_j b_j - T|^2 / (2 B) over samples X[B x 13] (10 cell features,
 * w_x, w_y, q) and labels T[B x 2], with c = MLP(10 -> h -> h -> 14) of the cell features;
 * gradients. terms: 7, or 3 (the quadrupole forms only; the other outputs get no gradient) */
double orc_learned_loss_grad(const double* X, const double* T, int64_t B, int h, int terms,
                             const double* W1, const double* b1, const double* W2,
                             const double* b2, const double* W3, const double* b3, double* gW1,
                             double* gb1, double* gW2, double* gb2, double* gW3, double* gb3);
/* gravity walk with single-precision decisions on given node records plus the correction on
 * accepted cells with d2 <= l2 * inv_band2 (single precision; inv_band2 = 0: all accepted
 * cells); coef: 14 doubles per node; counters = [pn, pp, corrected] */
void orc_bh_traverse_learned(const float* xs, const float* ys, const float* ms, int64_t n,
                             const int32_t* level, const uint32_t* first, const uint32_t* count,
                             const uint32_t* skip, const uint8_t* pp, const float* cx,
                             const float* cy, const float* mass, int64_t nnodes, float root_side,
                             float inv_theta2, double G, double eps2, const double* coef,
                             float inv_band2, const int64_t* sinks, int64_t nsinks,
                             double* out_ax, double* out_ay, uint64_t* counters);

#ifdef __cplusplus
}
#endif
#endif
