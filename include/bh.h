/*
 * bh.h — C ABI of the MI355X (gfx950) Barnes-Hut force step.
 *
 * Drop-in boundary for the reference's solver plug-in interface
 *   template <typename T> class NBodySolver            (src/nbodysolvers/cpu/nBodySolver.h:9-23)
 *     virtual void updateTree(std::vector<T>& points, glm::dvec2 minPos, glm::dvec2 maxPos);
 *     virtual void solveNbody(double& total, std::vector<T>& points);
 *     virtual std::vector<VertexPos2Col3> getNodesBufferData(int level);
 *     int maxChildren; double theta; double cell_size;
 * and of its GPU sibling NBodySolverGpu<T>::timeStep/getParticles
 *   (src/nbodysolvers/gpu/nBodySOlverGpu.cuh:9-33).
 * Plain C types only; no exceptions cross this boundary. A context owns all device memory,
 * never permutes caller memory, and returns results in the caller's particle order.
 * A context is not thread-safe; distinct contexts may be used from different threads.
 * Every call is synchronous on return unless its name ends in _async.
 */
#ifndef MTL_BH_H
#define MTL_BH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bh_ctx bh_ctx;

/* The reference reports no errors (void returns, CUDA errors unchecked,
 * cudaHelper.cuh:22-31); these status codes replace that. */
typedef enum bh_status {
    BH_OK = 0,
    BH_ERR_ARG = 1,      /* invalid argument (null pointer, n > n_max, bad config) */
    BH_ERR_HIP = 2,      /* HIP runtime error; see bh_last_error */
    BH_ERR_OOM = 3,      /* device allocation failed */
    BH_ERR_CAPACITY = 4, /* a fixed capacity was exceeded */
    BH_ERR_STATE = 5     /* call out of order (e.g. forces before a build) */
} bh_status;

/* interaction kernels: the reference injects them as std::function (nBodySolverBH.h:20-31) */
typedef enum bh_kernel {
    BH_KERNEL_GRAVITY = 0, /* a = G sum m_j (x_j - x)/(|x_j - x|^2 + eps^2)^(3/2) */
    BH_KERNEL_TSNE = 1     /* TSNEBHPNKernel / TSNEBHPPKernel (nBodySolverBH.h:94-116) */
} bh_kernel;

typedef enum bh_tree_mode {
    /* QuadTree (quadtree.h:29-171) + traverseBH (nBodySolverBH.h:56-89): adaptive, split while
     * > max_children, root square from the min corner with side max extent + 1e-4, a node with
     * at most one child is summed particle by particle over all its occupants. */
    BH_TREE_QUADTREE = 0,
    /* NBodySolverFMM_SYM_MORTON (nBodySolverFMM_SYM_MORTON.h:79-151, 272-313, 328-561):
     * 32-bit Morton codes, fixed depth getDepth(max_children, N), point-node traversal
     * accepting d^2 > (l/theta)^2, particle-particle at the leaves. */
    BH_TREE_MORTON_FIXED = 1,
    /* The tree of the reference's GPU step (NBodySolverGpuBH, nBodySolverGpuBH.cu:78-240):
     * complete to a fixed depth (config fixed_depth, 6 when 0, as tsneGpu.h:145 builds it),
     * root centred on the bounding box with side the larger extent (FP32), every cell split at
     * its centre with a coordinate equal to the centre going low, particle-particle at the
     * leaves. The acceptance is d^2 > (l/theta)^2 (the reference evaluates l/d < 1 with theta
     * fixed at 1: the two differ only when d and l agree to the last bit), and a leaf excludes
     * the sink itself by index (the reference compares positions, which also drops points
     * coincident with the sink). */
    BH_TREE_GPU_REF = 2
} bh_tree_mode;

typedef struct bh_config {
    uint32_t n_max;       /* particle capacity of the context */
    int32_t kernel;       /* bh_kernel */
    int32_t tree_mode;    /* bh_tree_mode */
    int32_t max_children; /* NBodySolver::maxChildren, leaf capacity (reference: 16, tsne.h:104) */
    int32_t fixed_depth;  /* MORTON_FIXED depth; 0 = getDepth(max_children, n) per build */
    int32_t code_bits;    /* QUADTREE: Morton bits per axis, 16..32 (0 = 32) */
    float G;              /* gravitational constant */
    float softening;      /* Plummer softening eps */
    double theta;         /* NBodySolver::theta, default opening angle */
} bh_config;

/* Fill cfg with the reference defaults: QuadTree mode, max_children 16, theta 0.5,
 * t-SNE kernel, 32 code bits. */
bh_status bh_config_default(bh_config* cfg);

/* Construct / destroy a solver (the reference's `new NBodySolverBH<...>(...)`, tsne.h:104-122;
 * device buffers are sized for cfg->n_max like NBodySolverGpuBH's constructor,
 * nBodySolverGpuBH.cu:489-555). */
bh_status bh_create(bh_ctx** out, const bh_config* cfg);
void bh_destroy(bh_ctx* ctx);
const char* bh_last_error(const bh_ctx* ctx);

/* Many independent systems in one context (no counterpart in the reference, whose instances
 * hold one system each): `systems` (<= 256) systems of sizes[s] > 0 particles, stored one after
 * the other in the order given. Each system gets its own QuadTree (root square, leaf capacity,
 * near-field rule and acceptance test as in BH_TREE_QUADTREE, with 24 code bits per axis: trees
 * deeper than 24 levels stop there); all systems are built by the same launches and traversed
 * by one launch whose waves each serve one system. Interactions never cross systems. Every
 * entry point below works on a batched context; bh_direct_fp64 sums over the sink's own
 * system. cfg->n_max is ignored (the total size is used). */
bh_status bh_create_batched(bh_ctx** out, const bh_config* cfg, const uint32_t* sizes,
                            uint32_t systems);
/* number of systems (0 for a one-system context) and, if offsets is not NULL, the
 * systems + 1 particle offsets */
bh_status bh_batch_info(bh_ctx* ctx, uint32_t* systems, uint32_t* offsets);

/* Upload particles in the caller's order (host pointers). m, vx, vy may be NULL:
 * masses default to 1 (the reference trees count particles, quadtree.h:163), velocities to 0.
 * The copy of a TsnePoint2D vector's positions that updateTree sees (tsne.h:202). */
bh_status bh_set_particles(bh_ctx* ctx, uint32_t n, const float* x, const float* y,
                           const float* m, const float* vx, const float* vy);
/* same with device pointers (inputs already resident in HBM) */
bh_status bh_set_particles_device(bh_ctx* ctx, uint32_t n, const float* x, const float* y,
                                  const float* m, const float* vx, const float* vy);
/* particle state in the caller's order (host pointers; any may be NULL) */
bh_status bh_get_particles(bh_ctx* ctx, float* x, float* y, float* vx, float* vy);

/* NBodySolver::updateTree(points, minPos, maxPos) (nBodySolver.h:18). bounds = {minx, miny,
 * maxx, maxy} as passed by the caller (tsne.h:260-270) or NULL to compute them on the device.
 * MORTON_FIXED honours caller bounds (nBodySolverFMM_SYM_MORTON.h:79-95); QUADTREE ignores
 * them like NBodySolverBH::updateTree (nBodySolverBH.h:43-46). */
bh_status bh_build(bh_ctx* ctx, const double* bounds);

/* NBodySolver::solveNbody(total, points) (nBodySolver.h:17, nBodySolverBH.h:33-41):
 * forces for the last built tree at opening angle theta. ax/ay (host, caller order, may be
 * NULL) receive the interaction sums (t-SNE: the unnormalised repulsive derivative; gravity:
 * the acceleration). total receives the FP64 sum of the per-particle totals (t-SNE: Z;
 * gravity: sum of particle potentials). */
bh_status bh_forces(bh_ctx* ctx, double theta, float* ax, float* ay, double* total);
/* same with device output pointers; total may be NULL */
bh_status bh_forces_device(bh_ctx* ctx, double theta, float* d_ax, float* d_ay, double* total);

/* Leapfrog kick-drift-kick for `steps` steps, rebuilding the tree every step (the order of
 * the reference's time step: update points, rebuild, solve; tsne_buffers.h:125-168).
 * Positions and velocities stay on the device; read them with bh_get_particles. */
bh_status bh_leapfrog(bh_ctx* ctx, int steps, float dt, double theta);
/* bh_leapfrog with each step (kick-drift, build, traversal, kick) captured once in a HIP graph
 * and replayed: same kernels and results, without per-launch host overhead. The graphs are
 * re-captured when dt, theta or the traversal variant change. */
bh_status bh_leapfrog_graph(bh_ctx* ctx, int steps, float dt, double theta);

/* Conserved quantities of a gravity context in FP64 (no counterpart in the reference, whose
 * gravity integrator is commented out, gravitysim.h:99-121): kinetic energy sum m |v|^2 / 2,
 * potential energy sum m phi / 2 with phi the per-particle potential of the Barnes-Hut walk at
 * theta on the current tree (G included), and momentum sum m v. Runs one walk with totals on the
 * last built tree (bh_leapfrog leaves one built at the current positions); the next bh_leapfrog
 * then evaluates its first forces anew. Any output may be NULL. */
bh_status bh_energy(bh_ctx* ctx, double theta, double* kinetic, double* potential,
                    double* momentum_x, double* momentum_y);

/* FP64 direct sum for sampled sinks (NBodySolverNaive::solveNbody with the same kernel,
 * nBodySolverNaive.h:53-140), over all current particles with the sink itself excluded by
 * index. sinks: caller-order indices (host). Outputs (host, FP64, nsinks each): the exact
 * interaction sums and per-sink totals. Used to report the Barnes-Hut error of a result. */
bh_status bh_direct_fp64(bh_ctx* ctx, const uint32_t* sinks, uint32_t nsinks, double* ax,
                         double* ay, double* total);

/* bh_direct_fp64 without any device allocation (no counterpart in the reference; for callers
 * that cannot allocate: code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). bh_direct_fp64 grows its sink and partial-sum buffers with hipMalloc on demand;
 * this variant keeps them in a static device array of the library (7.75 MiB per GPU, held until
 * the process ends) and so takes at most 65,536 sinks (BH_ERR_CAPACITY beyond). d_sinks
 * (caller-order indices) and the outputs (FP64, nsinks each; any output may be NULL) are device
 * pointers; an index >= the particle count gives NaN for that sink. Asynchronous: the work is
 * queued on the context's stream (bh_synchronize waits for it); the results equal
 * bh_direct_fp64's bit for bit. Calls that overlap on one GPU share the array and race. */
bh_status bh_direct_fp64_noalloc_async(bh_ctx* ctx, const uint32_t* d_sinks, uint32_t nsinks,
                                       double* d_ax, double* d_ay, double* d_total);

/* ---- t-SNE (the reference TSNE driver, tsne.h:42-563) ---- */
typedef struct bh_tsne_params {
    double learning_rate;    /* tsne.h:90: N / 15 */
    double momentum_early;   /* tsne.h:186: 0.5 before momentum_switch */
    double momentum_late;    /* 0.8 after */
    double exag_early;       /* tsne.h:358: 16 (4 x early exaggeration 4) before exag_switch */
    double exag_late;        /* 4 */
    double theta_min;        /* thetaFunction (tsne.h:208-233) */
    double theta_max;
    int32_t momentum_switch; /* 250 */
    int32_t exag_switch;     /* 250 */
    int32_t schedule_iters;  /* 999: theta(t) = min + (max-min)(1/2 + cos(pi min(t/999, 1))/2) */
    int32_t pad;
    double repulsion_scale;  /* tsne.h:291: derivative = (repulsion_scale / Z) rep + attraction;
                                4 (0 is read as 4) */
} bh_tsne_params;

/* reference constants for n points, constant theta (theta_min = theta_max = theta) */
bh_status bh_tsne_params_default(bh_tsne_params* p, uint32_t n, double theta);
/* the constants of the reference's GPU step (NBodySolverGpuBH::timeStep, tsneGpu.h:145,
 * nBodySolverGpuBH.cu:424-460): learning rate 1000, momentum 0.2, no exaggeration, theta 1,
 * derivative = (1/Z) rep + attraction. Use with a BH_TREE_GPU_REF context. */
bh_status bh_tsne_params_gpu_reference(bh_tsne_params* p);
/* theta used at iteration t: theta_max at t = 0, the cosine schedule at t - 1 afterwards */
double bh_tsne_theta(const bh_tsne_params* p, int32_t iteration);
/* P matrix in CSR, caller order, rows sorted by (row, col) as setupEfficientPmatrix builds it
 * (tsne.h:462-520); values FP32. The context must use BH_KERNEL_TSNE. */
bh_status bh_tsne_set_p(bh_ctx* ctx, uint32_t n, const uint32_t* row_ptr, const uint32_t* col,
                        const float* val, uint32_t nnz);
/* embedding (host, caller order); the previous positions start equal to it (tsne.h:553-555) */
bh_status bh_tsne_set_embedding(bh_ctx* ctx, const float* yx, const float* yy);
bh_status bh_tsne_get_embedding(bh_ctx* ctx, float* yx, float* yy);
/* one iteration t = updateDerivative + updatePoints (tsne.h:179-206, 272-301, 356-381):
 * theta(t) is theta_max at t = 0 and the schedule at t-1 afterwards, like the reference, whose
 * updatePoints sets the theta the next iteration uses. dx/dy (host, may be NULL) receive the
 * derivative d = (4/Z) repulsion + attraction at the pre-step embedding; z the Barnes-Hut Z. */
bh_status bh_tsne_step(bh_ctx* ctx, const bh_tsne_params* p, int32_t iteration, float* dx,
                       float* dy, double* z);
/* iterations first .. first+count-1 without host synchronisation in between */
bh_status bh_tsne_run(bh_ctx* ctx, const bh_tsne_params* p, int32_t first, int32_t count);
/* Input affinities of t-SNE from points X[n x dim] (FP32, row-major, dim <= 1024): exact k
 * nearest neighbours by Euclidean distance (self excluded, ties by index; FP16 MFMA candidates
 * with rigorous rounding bounds, FP64 distances decide), per-point Gaussian bandwidths beta_i
 * bisected until the entropy (natural log) of p_{j|i} = exp(-beta_i (d2_ij - d2_i0)) / sum is
 * within 1e-5 of log(perplexity) (at most 200 steps), and P = (P_cond + P_cond^T) / (2n) as CSR
 * sorted by (row, column): row_ptr[n + 1], col and val with room for 2 n k entries, *nnz set.
 * The reference computes P offline with openTSNE (MnistLoader.ipynb:539-547) and reads it as a
 * Matrix Market file (tsne.h:94-98, setupEfficientPmatrix tsne.h:462-520). Optional outputs:
 * knn_idx[n x k], knn_d2[n x k] (ascending), beta[n]. 1 <= k <= 128, k < n. Returns
 * BH_ERR_CAPACITY when some point has more than 192 candidates within the FP16 error bound of
 * its k-th neighbour (hundreds of exact or near duplicates). */
bh_status bh_tsne_affinities(bh_ctx* ctx, const float* X, uint32_t n, uint32_t dim,
                             double perplexity, uint32_t k, uint32_t* row_ptr, uint32_t* col,
                             float* val, uint32_t* nnz, uint32_t* knn_idx, double* knn_d2,
                             double* beta);
/* KL divergence of the current embedding with the exact FP64 Z (costFunction,
 * tsne.h:383-460; O(N^2) on the GPU in FP64) */
bh_status bh_tsne_kl(bh_ctx* ctx, double* kl, double* z_exact);

/* ---- learned far-field correction ----------------------------------------------------------
 * No counterpart in the reference. Per cell and build, a 10 -> 64 -> 64 -> 14 SiLU MLP maps the
 * cell's features (normalised second and third moments, centre-of-mass offset in its box, size)
 * to 7 complex coefficients c_j; an accepted (sink, cell) pair at d = c - p, r = |d|, side l then
 * gets (M l^2 / r^4) sum_j c_j b_j(d / r, l / r) added to its monopole, b the quadrupole and
 * octupole angular forms (exact definitions in masterthesislucas_amd/csrc/bh_mlp.h). Training
 * data come from the tree itself, labelled by FP64 direct sums over cell contents; training
 * runs on the GPU (FP32 forward/backward kernels, Adam); the per-cell inference uses BF16 MFMA
 * with FP32 accumulation. Gravity kernel, one-system BH_TREE_QUADTREE contexts only. */
typedef struct bh_mlp bh_mlp;
/* 5774: W1[10x64] b1 W2[64x64] b2 W3[64x14] b3, W row-major [in x out] */
uint32_t bh_mlp_param_count(void);
/* seeded uniform(+-1/sqrt(fan_in)) weights, zero biases; room for max_samples training pairs */
bh_status bh_mlp_create(bh_mlp** out, uint32_t max_samples, uint64_t seed);
void bh_mlp_destroy(bh_mlp* mlp);
const char* bh_mlp_last_error(const bh_mlp* mlp);
bh_status bh_mlp_set_params(bh_mlp* mlp, const float* params, uint32_t count); /* resets Adam */
bh_status bh_mlp_get_params(bh_mlp* mlp, float* params, uint32_t count);
/* basis terms of the loss and of the walk's correction: 7 (default, quadrupole and octupole
 * forms) or 3 (c_1 w + c_2 conj w + c_3 w^3 only: about half the correction's arithmetic);
 * set before training, the walk uses the network's setting */
bh_status bh_mlp_set_terms(bh_mlp* mlp, int32_t terms);
int32_t bh_mlp_get_terms(const bh_mlp* mlp);
/* coefficients Y[n x 14] of host cell features X[n x 10]: FP32 kernels (bf16_mfma = 0) or the
 * BF16 MFMA inference path the walk uses */
bh_status bh_mlp_forward(bh_mlp* mlp, const float* X, uint32_t n, float* Y, int bf16_mfma);
/* Samples X[n x 13] (cell features, w_x, w_y, q = l / r) with labels T[n x 2] (the residual
 * of the exact cell sum over the monopole, in units of M l^2 / r^4): loss = sum over samples of
 * |sum_j c_j b_j - T|^2 / (2 n) and its gradient (FP32 kernels) */
bh_status bh_mlp_loss_grad(bh_mlp* mlp, const float* X, const float* T, uint32_t n, float* grads,
                           double* loss);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected) with a given gradient */
bh_status bh_mlp_adam_step(bh_mlp* mlp, const float* grads, float lr);
bh_status bh_mlp_dataset_clear(bh_mlp* mlp);
bh_status bh_mlp_dataset_size(bh_mlp* mlp, uint32_t* n);
/* the data set: X[n x 13], T[n x 2] as in bh_mlp_loss_grad */
bh_status bh_mlp_dataset_get(bh_mlp* mlp, float* X, float* T, uint32_t n);
bh_status bh_mlp_dataset_set(bh_mlp* mlp, const float* X, const float* T, uint32_t n);
/* epochs over the data set in shuffled full batches; final_loss = the loss over the data set */
bh_status bh_mlp_train(bh_mlp* mlp, int epochs, uint32_t batch, float lr, uint64_t seed,
                       double* final_loss);
/* keep cell second moments and features in the builds of ctx (needed by the calls below) */
bh_status bh_learned_enable(bh_ctx* ctx, int enable);
/* append the (sink, cell) pairs the walk at theta accepts with l/r >= band (band 0: all) for
 * the given caller-order sinks, labelled by FP64 sums over the cells, to mlp's data set */
bh_status bh_learned_collect(bh_ctx* ctx, bh_mlp* mlp, double theta, double band,
                             const uint32_t* sinks, uint32_t nsinks, uint32_t max_pairs_per_sink,
                             uint32_t* added);
/* bh_forces with the learned correction on accepted cells with l/r >= band; the coefficient
 * records are computed from mlp for the current tree first (part of the call) */
bh_status bh_forces_learned(bh_ctx* ctx, const bh_mlp* mlp, double theta, double band, float* ax,
                            float* ay, double* total);
/* the current tree's cell features (features[n_nodes x 10]) and, with mlp, the walk's
 * coefficient records (coef[n_nodes x 16]: for c_j = a_j + i b_j, a1+a2, b2-b1, a1-a2, b1+b2,
 * a3, b3, a4, b4, a5+a6, b6-b5, a5-a6, b5+b6, a7, b7, then the cell's M l^2 and side l as the
 * walk uses them); n_nodes must equal the node count
 * of the tree (bh_tree_info). Either output may be NULL. */
bh_status bh_learned_export(bh_ctx* ctx, const bh_mlp* mlp, float* features, float* coef,
                            uint32_t n_nodes);

/* ---- failure reporting (no reference counterpart: the reference checks no errors,
 * cudaHelper.cuh:22-31) ----
 * Every call that builds a tree (bh_build, bh_leapfrog, bh_leapfrog_graph, bh_tsne_step,
 * bh_tsne_run, bh_tsne_kl) returns BH_ERR_ARG when a particle position is NaN or infinite (the
 * bounds pass flags it) and BH_ERR_HIP when the radix sort's look-back gave up waiting (the
 * order would not be valid); the tree is then marked unbuilt. Set the look-back bound (polls of
 * a tile waiting on an earlier tile; default 2^22, a wait that always ends on a working device)
 * to a small value only to exercise that error path. */
bh_status bh_set_sort_spin_limit(bh_ctx* ctx, uint32_t polls);

/* ---- introspection for parity checks (bit-exact against the reference tree) ---- */
/* bounds used by the last build {minx, miny, maxx, maxy} and the root side (FP64) */
bh_status bh_get_bounds(bh_ctx* ctx, double* bounds, double* root_side);
/* the quantiser the last build prepared on the device from the root side S (exact emulation of
 * floor(2^B * RN64(d / S)), masterthesislucas_amd/csrc/bh_softfp.h): S = M * 2^es and
 * R = floor((2^116 - 1) / M) */
bh_status bh_get_quantizer(bh_ctx* ctx, uint64_t* M, int32_t* es, uint64_t* R);
/* sorted Morton codes (createMortonCode, nBodySolverFMM_SYM_MORTON.h:538-552; 64-bit slots
 * holding 32-bit codes in MORTON_FIXED mode) and the sort permutation (original indices in
 * sorted order, integer_sort at :89 with ties by original index) */
bh_status bh_get_morton(bh_ctx* ctx, uint64_t* codes_sorted, uint32_t* perm);
/* node count, deepest leaf level and leaves at the depth limit holding > max_children */
bh_status bh_tree_info(bh_ctx* ctx, uint32_t* n_nodes, uint32_t* max_level,
                       uint32_t* overfull_leaves);
/* the pre-order tree: level, first sorted particle, count, skip, flags (bit0 leaf, bit1
 * particle-particle near field), centre of mass, mass; arrays of n_nodes (any may be NULL) */
bh_status bh_get_tree(bh_ctx* ctx, int32_t* level, uint32_t* first, uint32_t* count,
                      uint32_t* skip, uint8_t* flags, float* cx, float* cy, float* mass);
/* MORTON_FIXED only: the reference's slot arrays (firstChildIndex, firstParticleIndex,
 * particleIndexAmount of NodeFMM_MORTON_2D, nodeFMM_MORTON_2D.h:14-16) for all
 * (4^(depth+1)-1)/3 slots */
bh_status bh_get_tree_slots(bh_ctx* ctx, int32_t* depth, uint64_t* n_slots, uint32_t* first_child,
                            uint32_t* first_particle, uint32_t* count);
/* interaction counters of the last force evaluation: particle-node and particle-particle */
bh_status bh_counters(bh_ctx* ctx, uint64_t* pn, uint64_t* pp);

/* ---- timing (HIP events on the context's stream) ---- */
/* phases: 0 bounds+morton, 1 sort, 2 gather, 3 tree, 4 upward, 5 traversal, 6 integrate */
#define BH_NUM_PHASES 7
/* enable != 0 times every phase, 0 none. Each timed phase records two events per launch on the
 * stream, which delays the next kernel's start: timing every phase lengthens a 1M-particle
 * leapfrog step by about 5%. */
bh_status bh_profile_enable(bh_ctx* ctx, int enable);
/* time only the phases whose bit (1 << phase) is set in mask; BH_ERR_ARG for bits >= BH_NUM_PHASES */
bh_status bh_profile_phases(bh_ctx* ctx, uint32_t mask);
/* accumulated milliseconds and launch counts per phase since the last reset */
bh_status bh_profile_read(bh_ctx* ctx, double* ms, uint64_t* launches, int reset);
bh_status bh_synchronize(bh_ctx* ctx);

/* ---- autotuning: kernel variants of the traversal (no reference counterpart; the reference
 * GPU traversal has one fixed shape, nBodySolverGpuBH.cu:260-348) ----
 * variant v in [0, bh_traverse_variant_count()). For v < 12: v / 4 selects 1, 2 or 4 sink
 * particles per lane (a wave serves 64, 128 or 256 Morton-consecutive particles); (v % 4) / 2
 * selects monopoles computed only when a lane accepts (0) or always, masked (1); v % 2 selects
 * 1 or 4 sources per scalar fetch in the particle-particle loop. For 12 <= v < 18: (v - 12) / 2
 * selects 1, 2 or 4 sinks per lane, (v - 12) % 2 the monopole mode as above, and the
 * particle-particle sources of a cell are loaded by the lanes with one vector load (up to 64 at
 * a time) and broadcast. Variants 18 <= v < 36 take the shape of v - 18 with the lane masks
 * kept in scalar registers (acceptance mask = ballot of the distance test and the active mask;
 * fewer vector instructions per visit, faster for some shapes and workloads, slower for
 * others). Split walks (bh_set_traverse_split) use the vector-mask form of the shape. Results
 * are identical up to FP32 summation order. */
int bh_traverse_variant_count(void);
bh_status bh_set_traverse_variant(bh_ctx* ctx, int variant);
/* Split walk for small N (no reference counterpart): `split` waves (1..8) share each sink
 * group and take the same path through the tree; wave f adds the monopoles of every split-th
 * node visit and, in each particle-particle cell, the sources of the f-th of `split` equal
 * slices of the cell. Every interaction is evaluated by exactly one wave and the partial
 * results are summed in wave order. Multiplies the waves of a launch when the sinks alone
 * cannot fill the device and divides the long particle-particle loops of crowded leaves.
 * Results equal split 1 up to FP32 summation order; interaction counts are identical.
 * Batched contexts split each work entry the same way; bh_forces_learned always walks with
 * split 1. */
bh_status bh_set_traverse_split(bh_ctx* ctx, int split);
/* Issue priority of the walk's waves (no reference counterpart; on by default): each wave
 * starts at priority 3 and steps down a level every ~1/3 of a wave's node visits (a step set
 * from theta), so that waves that have made less progress are issued first and the waves of a
 * SIMD finish together instead of in dispatch order. Applied to one-system walks (not split,
 * not learned) whose waves all fit on the device at once (at most 8 per SIMD); other launches
 * walk without it. Scheduling only: results and interaction counts are bit-identical either
 * way. */
/* Batched contexts (no reference counterpart): rebuild the work table of the current variant
 * from the interaction counts of the last walk (split 1): entries in descending cost per lane
 * (a wave's interactions / its sinks x its sinks per lane), and the costliest entries (at least
 * 16 times the mean cost, at most 8192 of them) walked as 8 slices each (every slice takes the
 * path; slice f adds every 8th visit's monopoles and the f-th eighth of each particle-particle
 * cell's sources; the slices' partial results are added in slice order). Waves are otherwise
 * dispatched in system order, and on the 256-system workload the waves over the thousands of
 * occupants of one-child chains (up to 17 ms each) then start late: ordered, its walk takes
 * 13.4 ms instead of 21.8 (2 sinks per lane). Interaction counts are unchanged; forces change
 * only in the summation order of the sliced entries. Calling it again after a walk on the
 * ordered table gives the same table. bh_set_traverse_split(ctx, F > 1) restores the system
 * order. */
bh_status bh_batch_order_by_cost(bh_ctx* ctx);
bh_status bh_set_walk_priority(bh_ctx* ctx, int on);
bh_status bh_get_walk_priority(bh_ctx* ctx, int* on);

/* ---- force solver (NBodySolverFMM_SYM_MORTON, nBodySolverFMM_SYM_MORTON.h:28-316, 482-536,
 * registered as "FMM_SYM_MORTON" in tsne.h:121) ----
 * BH_SOLVER_BH: the Barnes-Hut walk (default). BH_SOLVER_FMM_SYM: the symmetric dual-tree FMM
 * with the t-SNE kernels (node-node expansions to third order with the source's second moment,
 * point-node, point-point; downward pass), in FP64, over the fixed-depth Morton tree. It needs a
 * one-system context with BH_KERNEL_TSNE and BH_TREE_MORTON_FIXED; bh_forces, bh_forces_device
 * and the bh_tsne_* steps then use it (theta > 0). Interaction counters read 0 for it. */
#define BH_SOLVER_BH 0
#define BH_SOLVER_FMM_SYM 1
bh_status bh_set_solver(bh_ctx* ctx, int solver);

#ifdef __cplusplus
}
#endif
#endif /* MTL_BH_H */
