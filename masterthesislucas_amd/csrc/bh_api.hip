// C ABI of the Barnes-Hut force step (declarations and reference citations in include/bh.h).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/bh.h"
#include "bh_common.h"
#include "bh_launch.h"
#include "bh_mlp.h"

using namespace bh;

namespace {

struct PhaseTimer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> spare;
    double ms = 0.0;
    uint64_t launches = 0;
};

// getDepth (nBodySolverFMM_SYM_MORTON.h:106-115), N/max_children an int division; 0 when
// N < max_children (the reference casts -inf to unsigned there).
int reference_depth(int max_children, uint32_t n) {
    const int q = (int)n / max_children;
    if (q <= 0) return 0;
    const double circle_area = M_PI / 4.0;
    const double depth_real = std::log((1.0 / circle_area) * (double)q) / std::log(4.0);
    const double c = std::ceil(depth_real);
    return c < 0 ? 0 : (int)c;
}

}  // namespace

struct bh_ctx {
    bh_config cfg{};
    std::string err;
    hipStream_t stream = nullptr;
    int device = 0;
    uint32_t n = 0;
    int cur = 0;
    bool built = false;
    bool moments = false;   // node records hold the upward pass of the last build
    bool have_acc = false;  // ax/ay valid for the current order (leapfrog)
    bool wide = true;       // 64-bit codes
    int B = 32;             // code bits per axis
    int depth = 0;          // MORTON_FIXED depth of the last build
    bool keys_in_b = false;
    uint32_t cap_nodes = 0;
    int max_fixed_depth = 0;

    float *x[2]{}, *y[2]{}, *m[2]{}, *vx[2]{}, *vy[2]{};
    uint32_t* id[2]{};
    float4* pos4 = nullptr;
    float *ax = nullptr, *ay = nullptr, *tot = nullptr;
    float *out_a = nullptr, *out_b = nullptr, *out_c = nullptr, *out_d = nullptr;
    void *keys_a = nullptr, *keys_b = nullptr;
    uint32_t *vals_a = nullptr, *vals_b = nullptr;
    uint32_t* sort_scratch = nullptr;  // radix sort counts, look-back words, totals, tickets
    uint32_t* scan_part = nullptr;
    uint8_t *leaf_lvl = nullptr, *lcp1 = nullptr;
    uint32_t *cnt = nullptr, *noff = nullptr;
    Node* nodes = nullptr;
    uint32_t *parent = nullptr, *child4 = nullptr, *deferred = nullptr, *bylevel = nullptr;
    float4* m2 = nullptr;
    BuildParams* prm = nullptr;
    float4* bounds_part = nullptr;
    double* red_part = nullptr;
    uint2* stats_part = nullptr;
    uint32_t max_blocks = 0;  // traversal blocks for n_max (counter pairs: one per wave)
    double* dsum = nullptr;
    unsigned long long* counters = nullptr;
    uint32_t *slot_a = nullptr, *slot_b = nullptr, *slot_c = nullptr;
    uint64_t slot_cap = 0;
    uint32_t* where = nullptr;           // inverse of the sorted-order ids
    uint32_t* dsinks = nullptr;          // direct sum: sink indices (caller, sorted)
    double* dpart = nullptr;             // direct sum: per-split partial sums + output
    uint64_t dsink_cap = 0, dpart_cap = 0;
    void* static_scratch = nullptr;      // this device's copy of the allocation-free direct
                                         // sum's static array (direct_static_scratch)

    uint32_t* dstatus = nullptr;   // build status words (kStatusWords), see status_check
    uint32_t* h_status = nullptr;  // pinned copy
    uint32_t sort_spin = kSortSpinLimit;
    BuildParams* h_prm = nullptr;  // pinned
    double* h_dsum = nullptr;
    unsigned long long* h_counters = nullptr;  // 2 per traversal wave
    uint2* h_stats = nullptr;

    // t-SNE state (caller order): P in CSR, embedding triple buffer, repulsion, derivative
    uint32_t *p_row = nullptr, *p_col = nullptr;
    float* p_val = nullptr;
    uint32_t p_n = 0, p_nnz = 0;
    float *ty_x[3]{}, *ty_y[3]{};
    float2* ty_xy = nullptr;  // interleaved copy of the current embedding (staged per iteration)
    int ty_cur = 0;  // ty[ty_cur] current, ty[(ty_cur+2)%3] previous
    bool ty_set = false;
    float *rep_x = nullptr, *rep_y = nullptr, *tder_x = nullptr, *tder_y = nullptr;
    float *att_x = nullptr, *att_y = nullptr;   // attraction sums (caller order)
    hipStream_t att_stream = nullptr;           // the attraction's stream (default priority)
    hipEvent_t ev_staged = nullptr, ev_att = nullptr;
    double* tz = nullptr;
    double* kl_part = nullptr;
    double* kl_out = nullptr;
    uint32_t* iota_dev = nullptr;

    bool learned = false;  // second moments in builds, learned correction available
    float* lfeat = nullptr;   // cell features of the last build (kMlpIn per node)
    float4* lcoef = nullptr;  // coefficient records (kCoefStride floats per node)
    uint32_t* lbig = nullptr; // large cells listed by the feature pass (counts + lists)
    bool lfeat_ok = false;    // lfeat holds the features of the current tree

    // batched systems (bh_create_batched): particles grouped by system, one tree per system
    uint32_t nsys = 0;             // 0: one system
    int R = 0;                     // level of the system roots in the tree (system id above)
    int qbits = 0;                 // quantisation bits per axis of batched codes
    std::vector<uint32_t> h_sys_off;
    uint32_t* sys_off = nullptr;   // nsys + 1 offsets
    BuildParams* sys_prm = nullptr;
    uint32_t* roots = nullptr;
    uint2* bblk = nullptr;         // bounds blocks: (first, count)
    uint32_t nbblk = 0;
    uint32_t* bblk_off = nullptr;  // first bounds block of each system, nsys + 1
    float4* bpart = nullptr;
    uint4* work[3]{};              // traversal work tables for 1, 2, 4 sinks per lane
    uint32_t n_work[3]{};
    std::vector<uint4> h_work0[3];  // the tables as built (system order, no heavy slices)
    uint4* heavy[3]{};              // per table: (first sink, count, slices) of its heavy slots
    uint32_t nheavy[3]{};
    float* side = nullptr;          // partial results of heavy slices (TraverseBatch::side)
    uint32_t side_cap = 0;          // heavy entries `side` holds room for

    // captured leapfrog steps (bh_leapfrog_graph), one per ping-pong parity
    hipGraph_t graph[2]{};
    hipGraphExec_t gexec[2]{};
    bool g_valid = false;
    float g_dt = 0.f;
    double g_theta = 0.0;
    int g_variant = -1;

    int variant = 4;  // traversal kernel variant (see launch_traverse)
    bool walk_prio = true;  // waves step their issue priority down as they progress
    uint32_t split = 1;         // waves per sink group in the traversal (one-system contexts)
    int solver = BH_SOLVER_BH;  // force solver (BH_SOLVER_FMM_SYM: symmetric dual-tree FMM)
    void* fmm_buf = nullptr;    // FMM workspace (fmm_workspace_bytes: per-slot moments,
                                // coefficients, interaction lists; per-point shares)
    float* part = nullptr;      // split traversal outputs: x, y, total parts, F n_max each
    uint32_t part_split = 0;    // F the part buffer was sized for
    uint32_t trav_blocks = 0;  // blocks of the last traversal launch
    uint32_t profile = 0;  // bit p: phase p is timed with HIP events
    PhaseTimer phases[BH_NUM_PHASES];

    std::vector<void*> allocs;
};

namespace {

bh_status fail(bh_ctx* c, bh_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return st;
}

bh_status hip_check(bh_ctx* c, hipError_t e, const char* what) {
    if (e == hipSuccess) return BH_OK;
    return fail(c, e == hipErrorOutOfMemory ? BH_ERR_OOM : BH_ERR_HIP, "%s: %s", what,
                hipGetErrorString(e));
}

template <typename T>
bh_status dalloc(bh_ctx* c, T** p, size_t count) {
    void* q = nullptr;
    const size_t bytes = count * sizeof(T) ? count * sizeof(T) : 16;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return hip_check(c, e, "hipMalloc");
    c->allocs.push_back(q);
    *p = (T*)q;
    return BH_OK;
}

void phase_begin(bh_ctx* c, int ph) {
    if (!((c->profile >> ph) & 1u)) return;
    PhaseTimer& t = c->phases[ph];
    std::pair<hipEvent_t, hipEvent_t> ev;
    if (!t.spare.empty()) {
        ev = t.spare.back();
        t.spare.pop_back();
    } else {
        (void)hipEventCreate(&ev.first);
        (void)hipEventCreate(&ev.second);
    }
    (void)hipEventRecord(ev.first, c->stream);
    t.pending.push_back(ev);
}

void phase_end(bh_ctx* c, int ph) {
    if (!((c->profile >> ph) & 1u)) return;
    PhaseTimer& t = c->phases[ph];
    (void)hipEventRecord(t.pending.back().second, c->stream);
    t.launches++;
}

void phase_collect(bh_ctx* c) {
    for (int p = 0; p < BH_NUM_PHASES; ++p) {
        PhaseTimer& t = c->phases[p];
        for (auto& ev : t.pending) {
            float ms = 0.f;
            if (hipEventSynchronize(ev.second) == hipSuccess &&
                hipEventElapsedTime(&ms, ev.first, ev.second) == hipSuccess)
                t.ms += ms;
            t.spare.push_back(ev);
        }
        t.pending.clear();
    }
}

bh_status launch_status(bh_ctx* c, const char* what) {
    return hip_check(c, hipGetLastError(), what);
}

// Build status: the words are cleared before a call's first build and copied to the host
// after its last launch; status_check reads them after the call's stream synchronisation.
bh_status status_begin(bh_ctx* c) {
    return hip_check(c, hipMemsetAsync(c->dstatus, 0, kStatusWords * sizeof(uint32_t), c->stream),
                     "status clear");
}

bh_status status_end(bh_ctx* c) {
    return hip_check(c, hipMemcpyAsync(c->h_status, c->dstatus, kStatusWords * sizeof(uint32_t),
                                       hipMemcpyDeviceToHost, c->stream),
                     "status copy");
}

// after the stream synchronisation that follows status_end
bh_status status_check(bh_ctx* c) {
    if (c->h_status[1]) {
        c->built = false;
        c->have_acc = false;
        return fail(c, BH_ERR_ARG, "a particle position is not finite (NaN or infinity)");
    }
    if (c->h_status[0]) {
        c->built = false;
        c->have_acc = false;
        return fail(c, BH_ERR_HIP,
                    "radix sort: a look-back wait exceeded %u polls; the sort order is not valid",
                    c->sort_spin);
    }
    return BH_OK;
}

// end of a call that built: status copy, synchronisation, status check
bh_status finish_build_call(bh_ctx* c, const char* what) {
    bh_status st = status_end(c);
    if (st != BH_OK) return st;
    st = hip_check(c, hipStreamSynchronize(c->stream), what);
    if (st != BH_OK) return st;
    return status_check(c);
}

// the reference GPU step's tree depth (NBodySolverGpuBH constructed with 6, tsneGpu.h:145)
constexpr int kGpuRefDepth = 6;

int effective_depth(const bh_ctx* c) {
    if (c->cfg.fixed_depth > 0) return c->cfg.fixed_depth;
    if (c->cfg.tree_mode == BH_TREE_GPU_REF) return kGpuRefDepth;
    return reference_depth(c->cfg.max_children, c->n);
}

void launch_node_moments(bh_ctx* c);

bh_status do_build(bh_ctx* c, const double* bounds) {
    const uint32_t n = c->n;
    if (n == 0) return fail(c, BH_ERR_STATE, "no particles");
    hipStream_t s = c->stream;
    const bool quadtree = c->cfg.tree_mode == BH_TREE_QUADTREE;
    const int cur = c->cur;
    int fixed = -1;
    if (!quadtree) {
        fixed = effective_depth(c);
        if (fixed > 16) return fail(c, BH_ERR_ARG, "fixed depth %d exceeds 16 (32-bit codes)", fixed);
        if (fixed > c->max_fixed_depth)
            return fail(c, BH_ERR_CAPACITY, "fixed depth %d above the capacity planned (%d)", fixed,
                        c->max_fixed_depth);
    }

    phase_begin(c, 0);
    if (c->nsys) {
        launch_bounds_batched(s, c->x[cur], c->y[cur], c->bblk, c->nbblk, c->bblk_off, c->sys_off,
                              c->nsys, c->bpart, c->sys_prm, c->prm, c->dstatus);
        launch_morton_batched(s, c->x[cur], c->y[cur], n, c->sys_off, c->nsys, c->sys_prm,
                              c->qbits, (uint64_t*)c->keys_a, c->vals_a);
    } else {
        if (c->cfg.tree_mode == BH_TREE_GPU_REF) {
            // the reference GPU step takes the bounds of the points themselves
            launch_bounds(s, c->x[cur], c->y[cur], n, c->bounds_part, c->prm, kTreeGpuRef, nullptr,
                          c->dstatus);
            launch_morton_gpuref(s, c->x[cur], c->y[cur], n, c->prm, fixed, (uint32_t*)c->keys_a,
                                 c->vals_a);
        } else {
            launch_bounds(s, c->x[cur], c->y[cur], n, c->bounds_part, c->prm,
                          quadtree ? kTreeQuadTree : kTreeMortonFixed, quadtree ? nullptr : bounds,
                          c->dstatus);
            launch_morton(s, c->x[cur], c->y[cur], n, c->prm, c->B, c->keys_a, c->vals_a,
                          c->wide);
        }
    }
    phase_end(c, 0);

    phase_begin(c, 1);
    bool in_b = false;
    const int key_bits = c->nsys ? 2 * c->qbits + 2 * c->R : 2 * c->B;
    launch_radix_sort(s, c->keys_a, c->vals_a, c->keys_b, c->vals_b, n, key_bits, c->wide,
                      c->sort_scratch, c->dstatus, c->sort_spin, &in_b);
    phase_end(c, 1);
    c->keys_in_b = in_b;
    void* codes = in_b ? c->keys_b : c->keys_a;
    const uint32_t* perm = in_b ? c->vals_b : c->vals_a;

    phase_begin(c, 2);
    const int nxt = 1 - cur;
    launch_gather(s, perm, n, c->x[cur], c->y[cur], c->m[cur], c->vx[cur], c->vy[cur], c->id[cur],
                  c->x[nxt], c->y[nxt], c->m[nxt], c->vx[nxt], c->vy[nxt], c->id[nxt], c->pos4,
                  c->dstatus, codes, c->wide);
    phase_end(c, 2);
    c->cur = nxt;

    if (!quadtree) c->depth = fixed;
    phase_begin(c, 3);
    launch_tree(s, codes, c->wide, n, c->B, c->cfg.max_children, fixed, quadtree ? 1 : 0,
                c->leaf_lvl, c->lcp1, c->cnt, c->noff, c->scan_part, c->prm, c->nodes, c->parent,
                c->child4, c->cap_nodes, c->R, c->nsys ? c->sys_prm : nullptr, c->dstatus);
    if (c->nsys) launch_batch_roots(s, c->sys_off, c->nsys, c->noff, c->prm, c->roots);
    phase_end(c, 3);

    // the FMM solver computes its own moments: the node records' moments are made when a
    // consumer of them asks (ensure_moments)
    c->moments = false;
    c->lfeat_ok = false;
    if (c->solver != BH_SOLVER_FMM_SYM) launch_node_moments(c);
    bh_status st = launch_status(c, "build launch");
    if (st != BH_OK) return st;
    c->built = true;
    c->have_acc = false;
    return BH_OK;
}

// upward pass over the node records (mass, centre of mass, second moments when learned)
void launch_node_moments(bh_ctx* c) {
    phase_begin(c, 4);
    launch_upward(c->stream, c->pos4, c->nodes, c->child4, c->learned ? c->m2 : nullptr, c->prm,
                  c->deferred, c->bylevel, c->cap_nodes, !c->wide && !c->nsys);
    phase_end(c, 4);
    c->moments = true;
}

bh_status ensure_moments(bh_ctx* c) {
    if (c->moments) return BH_OK;
    launch_node_moments(c);
    return launch_status(c, "upward launch");
}

TraverseBatch batch_of(const bh_ctx* c, int ns_index) {
    TraverseBatch b;
    if (c->nsys) {
        b.work = c->work[ns_index];
        b.n_work = c->n_work[ns_index];
        b.roots = c->roots;
        b.side = c->side;
        b.side_slots = c->side_cap;
    }
    return b;
}

bh_status do_fmm(bh_ctx* c, double theta) {
    if (!(theta > 0.0)) return fail(c, BH_ERR_ARG, "the FMM solver needs theta > 0");
    hipStream_t s = c->stream;
    phase_begin(c, 5);
    const uint32_t* keys = (const uint32_t*)(c->keys_in_b ? c->keys_b : c->keys_a);
    launch_tree_slots(s, keys, c->nodes, c->prm, c->cap_nodes, c->depth, c->slot_a, c->slot_b,
                      c->slot_c, fmm_slots(c->depth));
    launch_fmm(s, c->pos4, keys, c->slot_b, c->slot_c, c->depth, c->slot_cap, c->n, c->cfg.n_max,
               c->prm, theta, c->fmm_buf, c->ax, c->ay, c->dsum);
    phase_end(c, 5);
    c->trav_blocks = 0;  // interaction counters are those of the tree walk only
    c->have_acc = true;
    return launch_status(c, "FMM launch");
}

bh_status do_forces(bh_ctx* c, double theta, bool want_total) {
    if (!c->built) return fail(c, BH_ERR_STATE, "forces requested before a tree was built");
    if (!(theta >= 0.0)) return fail(c, BH_ERR_ARG, "theta must be >= 0");
    if (c->solver == BH_SOLVER_FMM_SYM) return do_fmm(c, theta);
    bh_status mst = ensure_moments(c);
    if (mst != BH_OK) return mst;
    hipStream_t s = c->stream;
    const float inv_theta2 = (float)(1.0 / (theta * theta));
    const float eps = c->cfg.softening;
    const bool grav = c->cfg.kernel == BH_KERNEL_GRAVITY;
    // a split walk writes F parts of n per output, summed in part order afterwards
    const uint32_t F = c->split;
    const size_t pn = (size_t)F * c->cfg.n_max;
    float* ox = F > 1 ? c->part : c->ax;
    float* oy = F > 1 ? c->part + pn : c->ay;
    float* ot = F > 1 ? c->part + 2 * pn : c->tot;
    phase_begin(c, 5);
    if (!c->nsys && (uint64_t)c->n * 16 <= (3ull << 20))  // small trees: warm every L2 first
        launch_warm_l2(s, c->nodes, c->prm, c->pos4, c->n, c->vals_b);
    c->trav_blocks = launch_traverse(s, grav ? kKernelGravity : kKernelTsne, c->variant, c->pos4,
                                     c->nodes, c->prm, c->n, inv_theta2, eps * eps,
                                     grav ? c->cfg.G : 1.0f, ox, oy, want_total ? ot : nullptr,
                                     c->counters,
                                     batch_of(c, variant_ns_index(c->variant)),
                                     F, c->walk_prio);
    if (F > 1) {
        launch_sum_parts(s, ox, F, c->n, c->ax);
        launch_sum_parts(s, oy, F, c->n, c->ay);
        if (want_total) launch_sum_parts(s, ot, F, c->n, c->tot);
    } else if (c->nsys) {
        const int v = variant_ns_index(c->variant);
        launch_sum_heavy(s, c->side, c->side_cap, c->heavy[v], c->nheavy[v], c->ax, c->ay,
                         want_total ? c->tot : nullptr);
    }
    phase_end(c, 5);
    if (want_total)
        launch_reduce_f64(s, c->tot, nullptr, c->n, c->red_part, c->dsum);
    c->have_acc = true;
    return launch_status(c, "traverse launch");
}

}  // namespace

extern "C" {

bh_status bh_config_default(bh_config* cfg) {
    if (!cfg) return BH_ERR_ARG;
    memset(cfg, 0, sizeof(*cfg));
    cfg->n_max = 0;
    cfg->kernel = BH_KERNEL_TSNE;
    cfg->tree_mode = BH_TREE_QUADTREE;
    cfg->max_children = 16;
    cfg->fixed_depth = 0;
    cfg->code_bits = 32;
    cfg->G = 1.0f;
    cfg->softening = 0.0f;
    cfg->theta = 0.5;
    return BH_OK;
}

const char* bh_last_error(const bh_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

static void graph_reset(bh_ctx* c);

void bh_destroy(bh_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->att_stream) {
        (void)hipStreamSynchronize(c->att_stream);
        (void)hipStreamDestroy(c->att_stream);
    }
    if (c->ev_staged) (void)hipEventDestroy(c->ev_staged);
    if (c->ev_att) (void)hipEventDestroy(c->ev_att);
    graph_reset(c);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->dsinks) (void)hipFree(c->dsinks);
    if (c->p_row) (void)hipFree(c->p_row);
    if (c->p_col) (void)hipFree(c->p_col);
    if (c->p_val) (void)hipFree(c->p_val);
    if (c->dpart) (void)hipFree(c->dpart);
    if (c->part) (void)hipFree(c->part);
    if (c->fmm_buf) (void)hipFree(c->fmm_buf);
    if (c->side) (void)hipFree(c->side);
    if (c->h_status) (void)hipHostFree(c->h_status);
    if (c->h_prm) (void)hipHostFree(c->h_prm);
    if (c->h_dsum) (void)hipHostFree(c->h_dsum);
    if (c->h_counters) (void)hipHostFree(c->h_counters);
    if (c->h_stats) (void)hipHostFree(c->h_stats);
    for (auto& t : c->phases) {
        for (auto& e : t.pending) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        for (auto& e : t.spare) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

static bh_status create_impl(bh_ctx** out, const bh_config* cfg, uint32_t extra_nodes,
                             uint32_t extra_blocks) {
    if (!out || !cfg) return BH_ERR_ARG;
    *out = nullptr;
    if (cfg->n_max == 0 || cfg->n_max > (1u << 30)) return BH_ERR_ARG;
    if (cfg->kernel != BH_KERNEL_GRAVITY && cfg->kernel != BH_KERNEL_TSNE) return BH_ERR_ARG;
    if (cfg->tree_mode != BH_TREE_QUADTREE && cfg->tree_mode != BH_TREE_MORTON_FIXED &&
        cfg->tree_mode != BH_TREE_GPU_REF)
        return BH_ERR_ARG;
    if (cfg->max_children < 1 || cfg->max_children > kMaxLeafCap) return BH_ERR_ARG;
    if (cfg->fixed_depth < 0 || cfg->fixed_depth > 16) return BH_ERR_ARG;
    const int bits = cfg->code_bits ? cfg->code_bits : 32;
    if (bits < 16 || bits > 32) return BH_ERR_ARG;

    bh_ctx* c = new (std::nothrow) bh_ctx();
    if (!c) return BH_ERR_OOM;
    c->cfg = *cfg;
    c->cfg.code_bits = bits;
    c->wide = cfg->tree_mode == BH_TREE_QUADTREE;
    c->B = c->wide ? bits : 16;

    bh_status st;
#define TRY(expr)                 \
    do {                          \
        st = (expr);              \
        if (st != BH_OK) {        \
            *out = c;             \
            return st;            \
        }                         \
    } while (0)
    TRY(hip_check(c, hipGetDevice(&c->device), "hipGetDevice"));
    TRY(hip_check(c, direct_static_scratch(c->device, &c->static_scratch),
                  "hipGetSymbolAddress (static direct-sum scratch)"));
    TRY(hip_check(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate"));
    const uint32_t N = cfg->n_max;
    if (c->wide) {
        // cells of one level are disjoint and an internal cell holds > max_children particles:
        // at most N leaves plus B * N / (max_children + 1) internal cells
        const uint64_t capn = (uint64_t)N + (uint64_t)c->B * N / (uint64_t)(cfg->max_children + 1) +
                              64 + extra_nodes;
        if (capn > 0xFFFFFFF0ull) {
            *out = c;
            return fail(c, BH_ERR_CAPACITY, "node capacity overflows 32 bits");
        }
        c->cap_nodes = (uint32_t)capn;
    } else {
        const int dcap = cfg->tree_mode == BH_TREE_GPU_REF
                             ? (cfg->fixed_depth > 0 ? cfg->fixed_depth : kGpuRefDepth)
                             : std::max(cfg->fixed_depth, reference_depth(cfg->max_children, N));
        c->max_fixed_depth = std::max(dcap, 1);
        const uint64_t slots = ((1ull << (2 * (c->max_fixed_depth + 1))) - 1) / 3;
        const uint64_t bound = (uint64_t)(c->max_fixed_depth + 1) * N + 1;
        c->cap_nodes = (uint32_t)std::min(slots, bound);
        c->slot_cap = slots;
    }
    for (int k = 0; k < 2; ++k) {
        TRY(dalloc(c, &c->x[k], N));
        TRY(dalloc(c, &c->y[k], N));
        TRY(dalloc(c, &c->m[k], N));
        TRY(dalloc(c, &c->vx[k], N));
        TRY(dalloc(c, &c->vy[k], N));
        TRY(dalloc(c, &c->id[k], N));
    }
    TRY(dalloc(c, &c->pos4, (size_t)N + kPosPad));
    TRY(dalloc(c, &c->ax, N));
    TRY(dalloc(c, &c->ay, N));
    TRY(dalloc(c, &c->tot, N));
    TRY(dalloc(c, &c->out_a, N));
    TRY(dalloc(c, &c->out_b, N));
    TRY(dalloc(c, &c->out_c, N));
    TRY(dalloc(c, &c->out_d, N));
    const size_t ksz = c->wide ? sizeof(uint64_t) : sizeof(uint32_t);
    {
        void* q;
        TRY(hip_check(c, hipMalloc(&q, (size_t)N * ksz), "hipMalloc keys"));
        c->allocs.push_back(q);
        c->keys_a = q;
        TRY(hip_check(c, hipMalloc(&q, (size_t)N * ksz), "hipMalloc keys"));
        c->allocs.push_back(q);
        c->keys_b = q;
    }
    TRY(dalloc(c, &c->vals_a, N));
    TRY(dalloc(c, &c->vals_b, N));
    TRY(dalloc(c, &c->sort_scratch, sort_scratch_words(N)));
    TRY(dalloc(c, &c->scan_part, scan_partials_size(N) + 256));
    TRY(dalloc(c, &c->leaf_lvl, N));
    TRY(dalloc(c, &c->lcp1, N));
    TRY(dalloc(c, &c->cnt, N));
    TRY(dalloc(c, &c->noff, N));
    TRY(dalloc(c, &c->nodes, c->cap_nodes));
    TRY(dalloc(c, &c->parent, c->cap_nodes));
    TRY(dalloc(c, &c->deferred, (size_t)up_chunks(c->cap_nodes) * kDeferStride));
    TRY(dalloc(c, &c->bylevel, (size_t)up_chunks(c->cap_nodes) * kDeferStride));
    TRY(dalloc(c, &c->stats_part, kStatsBlocks));
    TRY(dalloc(c, &c->child4, 4ull * c->cap_nodes));
    TRY(dalloc(c, &c->prm, 1));
    TRY(dalloc(c, &c->bounds_part, kMaxBoundsBlocks));
    TRY(dalloc(c, &c->red_part, kMaxReduceBlocks));
    TRY(dalloc(c, &c->dsum, 5));  // [0] the walk's FP64 total, [1..4] state sums
    c->max_blocks = traverse_blocks(N) * kMaxSplit + extra_blocks;
    TRY(dalloc(c, &c->counters, 2ull * kWavesPerBlock * c->max_blocks));
    TRY(dalloc(c, &c->where, N));
    if (cfg->kernel == BH_KERNEL_TSNE) {
        for (int k = 0; k < 3; ++k) {
            TRY(dalloc(c, &c->ty_x[k], N));
            TRY(dalloc(c, &c->ty_y[k], N));
        }
        TRY(dalloc(c, &c->ty_xy, N));
        TRY(dalloc(c, &c->rep_x, N));
        TRY(dalloc(c, &c->rep_y, N));
        TRY(dalloc(c, &c->tder_x, N));
        TRY(dalloc(c, &c->tder_y, N));
        TRY(dalloc(c, &c->att_x, N));
        TRY(dalloc(c, &c->att_y, N));
        TRY(hip_check(c, hipStreamCreateWithFlags(&c->att_stream, hipStreamNonBlocking),
                      "hipStreamCreate"));
        TRY(hip_check(c, hipEventCreateWithFlags(&c->ev_staged, hipEventDisableTiming),
                      "hipEventCreate"));
        TRY(hip_check(c, hipEventCreateWithFlags(&c->ev_att, hipEventDisableTiming),
                      "hipEventCreate"));
        TRY(dalloc(c, &c->tz, 2));
        TRY(dalloc(c, &c->kl_part, kMaxReduceBlocks));
        TRY(dalloc(c, &c->kl_out, 2));
        TRY(dalloc(c, &c->iota_dev, N));
        launch_iota(c->stream, c->iota_dev, N);
    }
    if (!c->wide) {
        TRY(dalloc(c, &c->slot_a, c->slot_cap));
        TRY(dalloc(c, &c->slot_b, c->slot_cap));
        TRY(dalloc(c, &c->slot_c, c->slot_cap));
    }
    TRY(dalloc(c, &c->dstatus, kStatusWords));
    TRY(hip_check(c, hipHostMalloc((void**)&c->h_status, kStatusWords * sizeof(uint32_t), 0),
                  "hipHostMalloc"));
    memset(c->h_status, 0, kStatusWords * sizeof(uint32_t));
    launch_fill_u32(c->stream, c->dstatus, kStatusWords, 0u);
    TRY(hip_check(c, hipHostMalloc((void**)&c->h_prm, sizeof(BuildParams), 0), "hipHostMalloc"));
    TRY(hip_check(c, hipHostMalloc((void**)&c->h_dsum, sizeof(double), 0), "hipHostMalloc"));
    TRY(hip_check(c, hipHostMalloc((void**)&c->h_stats, kStatsBlocks * sizeof(uint2), 0),
                  "hipHostMalloc"));
    TRY(hip_check(c, hipHostMalloc((void**)&c->h_counters,
                                   2ull * kWavesPerBlock * c->max_blocks * sizeof(unsigned long long), 0),
                  "hipHostMalloc"));
    launch_fill_f32(c->stream, (float*)c->pos4, 4ull * ((uint64_t)N + kPosPad), 0.0f);
    launch_fill_u32(c->stream, (uint32_t*)c->prm, sizeof(BuildParams) / 4, 0u);
    TRY(hip_check(c, hipStreamSynchronize(c->stream), "init"));
#undef TRY
    *out = c;
    return BH_OK;
}

bh_status bh_create(bh_ctx** out, const bh_config* cfg) { return create_impl(out, cfg, 0, 0); }

// batched systems: codes of kBatchBits per axis under an 8-bit system id (4 tree levels)
constexpr int kBatchBits = 24;
constexpr int kBatchRootLevel = 4;
constexpr uint32_t kBatchMaxSystems = 256;
constexpr uint32_t kBatchBoundsChunk = 8192;

bh_status bh_create_batched(bh_ctx** out, const bh_config* cfg, const uint32_t* sizes,
                            uint32_t systems) {
    if (!out || !cfg || !sizes) return BH_ERR_ARG;
    *out = nullptr;
    if (systems == 0 || systems > kBatchMaxSystems) return BH_ERR_ARG;
    if (cfg->tree_mode != BH_TREE_QUADTREE) return BH_ERR_ARG;
    uint64_t N = 0;
    for (uint32_t s = 0; s < systems; ++s) {
        if (sizes[s] == 0) return BH_ERR_ARG;
        N += sizes[s];
    }
    if (N > (1u << 30)) return BH_ERR_ARG;
    bh_config c2 = *cfg;
    c2.n_max = (uint32_t)N;
    c2.code_bits = kBatchBits + kBatchRootLevel;  // tree levels counted from the system id
    bh_status st = create_impl(out, &c2, 64u * systems, systems + 1);
    if (st != BH_OK) return st;
    bh_ctx* c = *out;
    c->nsys = systems;
    c->R = kBatchRootLevel;
    c->qbits = kBatchBits;
    c->h_sys_off.assign(systems + 1, 0);
    for (uint32_t s = 0; s < systems; ++s) c->h_sys_off[s + 1] = c->h_sys_off[s] + sizes[s];
    std::vector<uint2> blk;
    std::vector<uint32_t> blk_off(systems + 1, 0);
    for (uint32_t s = 0; s < systems; ++s) {
        for (uint32_t f = 0; f < sizes[s]; f += kBatchBoundsChunk)
            blk.push_back(make_uint2(c->h_sys_off[s] + f, std::min(kBatchBoundsChunk, sizes[s] - f)));
        blk_off[s + 1] = (uint32_t)blk.size();
    }
    c->nbblk = (uint32_t)blk.size();
#define TRYB(expr)                \
    do {                          \
        st = (expr);              \
        if (st != BH_OK) return st; \
    } while (0)
    TRYB(dalloc(c, &c->sys_off, systems + 1));
    TRYB(dalloc(c, &c->sys_prm, systems));
    TRYB(dalloc(c, &c->roots, systems));
    TRYB(dalloc(c, &c->bblk, c->nbblk));
    TRYB(dalloc(c, &c->bblk_off, systems + 1));
    TRYB(dalloc(c, &c->bpart, c->nbblk));
    hipStream_t s = c->stream;
    TRYB(hip_check(c, hipMemcpyAsync(c->sys_off, c->h_sys_off.data(), (systems + 1) * 4,
                                     hipMemcpyHostToDevice, s), "sys_off"));
    TRYB(hip_check(c, hipMemcpyAsync(c->bblk, blk.data(), blk.size() * sizeof(uint2),
                                     hipMemcpyHostToDevice, s), "blocks"));
    TRYB(hip_check(c, hipMemcpyAsync(c->bblk_off, blk_off.data(), blk_off.size() * 4,
                                     hipMemcpyHostToDevice, s), "blocks"));
    for (int v = 0; v < 3; ++v) {
        const uint32_t per = 64u << v;  // sinks per wave for 1, 2, 4 sinks per lane
        std::vector<uint4> wk;
        for (uint32_t q = 0; q < systems; ++q)
            for (uint32_t f = 0; f < sizes[q]; f += per)
                wk.push_back(make_uint4(c->h_sys_off[q] + f, std::min(per, sizes[q] - f), q, 0u));
        c->n_work[v] = (uint32_t)wk.size();
        // room for the heavy entries' extra slices (bh_batch_order_by_cost)
        TRYB(dalloc(c, &c->work[v], wk.size() + (size_t)kHeavyMax * (kHeavySlices - 1)));
        TRYB(dalloc(c, &c->heavy[v], kHeavyMax));
        TRYB(hip_check(c, hipMemcpyAsync(c->work[v], wk.data(), wk.size() * sizeof(uint4),
                                         hipMemcpyHostToDevice, s), "work"));
        c->h_work0[v] = std::move(wk);
    }
#undef TRYB
    return hip_check(c, hipStreamSynchronize(s), "batched init");
}

bh_status bh_batch_info(bh_ctx* c, uint32_t* systems, uint32_t* offsets) {
    if (!c || !systems) return BH_ERR_ARG;
    *systems = c->nsys;
    if (offsets)
        for (uint32_t s = 0; s <= c->nsys; ++s) offsets[s] = c->h_sys_off[s];
    return BH_OK;
}

static bh_status set_particles(bh_ctx* c, uint32_t n, const float* x, const float* y,
                               const float* m, const float* vx, const float* vy, bool device) {
    if (!c || !x || !y) return BH_ERR_ARG;
    if (n > c->cfg.n_max) return fail(c, BH_ERR_ARG, "n=%u exceeds n_max=%u", n, c->cfg.n_max);
    if (c->nsys && n != c->h_sys_off[c->nsys])
        return fail(c, BH_ERR_ARG, "batched systems hold %u particles, got %u", c->h_sys_off[c->nsys], n);
    if ((vx == nullptr) != (vy == nullptr)) return fail(c, BH_ERR_ARG, "vx and vy go together");
    hipStream_t s = c->stream;
    const hipMemcpyKind kind = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (device) {
        bh_status st = hip_check(c, hipDeviceSynchronize(), "sync before device upload");
        if (st != BH_OK) return st;
    }
    c->cur = 0;
    c->n = n;
    c->built = false;
    c->have_acc = false;
    if (n == 0) return BH_OK;
    const size_t b = (size_t)n * sizeof(float);
    bh_status st = hip_check(c, hipMemcpyAsync(c->x[0], x, b, kind, s), "upload x");
    if (st == BH_OK) st = hip_check(c, hipMemcpyAsync(c->y[0], y, b, kind, s), "upload y");
    if (st != BH_OK) return st;
    if (m)
        st = hip_check(c, hipMemcpyAsync(c->m[0], m, b, kind, s), "upload m");
    else
        launch_fill_f32(s, c->m[0], n, 1.0f);
    if (st != BH_OK) return st;
    if (vx) {
        st = hip_check(c, hipMemcpyAsync(c->vx[0], vx, b, kind, s), "upload vx");
        if (st == BH_OK) st = hip_check(c, hipMemcpyAsync(c->vy[0], vy, b, kind, s), "upload vy");
        if (st != BH_OK) return st;
    } else {
        launch_fill_f32(s, c->vx[0], n, 0.0f);
        launch_fill_f32(s, c->vy[0], n, 0.0f);
    }
    launch_iota(s, c->id[0], n);
    return hip_check(c, hipStreamSynchronize(s), "upload");
}

bh_status bh_set_particles(bh_ctx* c, uint32_t n, const float* x, const float* y, const float* m,
                           const float* vx, const float* vy) {
    return set_particles(c, n, x, y, m, vx, vy, false);
}

bh_status bh_set_particles_device(bh_ctx* c, uint32_t n, const float* x, const float* y,
                                  const float* m, const float* vx, const float* vy) {
    return set_particles(c, n, x, y, m, vx, vy, true);
}

bh_status bh_get_particles(bh_ctx* c, float* x, float* y, float* vx, float* vy) {
    if (!c) return BH_ERR_ARG;
    const uint32_t n = c->n;
    if (n == 0) return BH_OK;
    hipStream_t s = c->stream;
    const int cur = c->cur;
    launch_unsort2(s, c->id[cur], n, c->x[cur], c->y[cur], c->out_a, c->out_b);
    launch_unsort2(s, c->id[cur], n, c->vx[cur], c->vy[cur], c->out_c, c->out_d);
    const size_t b = (size_t)n * sizeof(float);
    bh_status st = BH_OK;
    if (x) st = hip_check(c, hipMemcpyAsync(x, c->out_a, b, hipMemcpyDeviceToHost, s), "x");
    if (st == BH_OK && y) st = hip_check(c, hipMemcpyAsync(y, c->out_b, b, hipMemcpyDeviceToHost, s), "y");
    if (st == BH_OK && vx) st = hip_check(c, hipMemcpyAsync(vx, c->out_c, b, hipMemcpyDeviceToHost, s), "vx");
    if (st == BH_OK && vy) st = hip_check(c, hipMemcpyAsync(vy, c->out_d, b, hipMemcpyDeviceToHost, s), "vy");
    if (st != BH_OK) return st;
    return hip_check(c, hipStreamSynchronize(s), "get particles");
}

bh_status bh_build(bh_ctx* c, const double* bounds) {
    if (!c) return BH_ERR_ARG;
    bh_status st = status_begin(c);
    if (st == BH_OK) st = do_build(c, bounds);
    if (st != BH_OK) return st;
    return finish_build_call(c, "build");
}

static bh_status forces_out(bh_ctx* c, double theta, float* ax, float* ay, double* total,
                            bool device) {
    if (!c) return BH_ERR_ARG;
    bh_status st = do_forces(c, theta, total != nullptr);
    if (st != BH_OK) return st;
    hipStream_t s = c->stream;
    const uint32_t n = c->n;
    const int cur = c->cur;
    const size_t b = (size_t)n * sizeof(float);
    if (ax || ay) {
        if (device) {
            launch_unsort2(s, c->id[cur], n, c->ax, c->ay, ax ? ax : c->out_a, ay ? ay : c->out_b);
        } else {
            launch_unsort2(s, c->id[cur], n, c->ax, c->ay, c->out_a, c->out_b);
            if (ax) st = hip_check(c, hipMemcpyAsync(ax, c->out_a, b, hipMemcpyDeviceToHost, s), "ax");
            if (st == BH_OK && ay)
                st = hip_check(c, hipMemcpyAsync(ay, c->out_b, b, hipMemcpyDeviceToHost, s), "ay");
            if (st != BH_OK) return st;
        }
    }
    if (total) {
        st = hip_check(c, hipMemcpyAsync(c->h_dsum, c->dsum, sizeof(double), hipMemcpyDeviceToHost, s),
                       "total");
        if (st != BH_OK) return st;
    }
    st = hip_check(c, hipStreamSynchronize(s), "forces");
    if (st != BH_OK) return st;
    if (total) *total = *c->h_dsum;
    return BH_OK;
}

bh_status bh_forces(bh_ctx* c, double theta, float* ax, float* ay, double* total) {
    return forces_out(c, theta, ax, ay, total, false);
}

bh_status bh_forces_device(bh_ctx* c, double theta, float* d_ax, float* d_ay, double* total) {
    return forces_out(c, theta, d_ax, d_ay, total, true);
}

bh_status bh_leapfrog(bh_ctx* c, int steps, float dt, double theta) {
    if (!c || steps < 0) return BH_ERR_ARG;
    if (c->n == 0) return fail(c, BH_ERR_STATE, "no particles");
    bh_status st = status_begin(c);
    if (st != BH_OK) return st;
    if (!c->have_acc) {
        st = do_build(c, nullptr);
        if (st == BH_OK) st = do_forces(c, theta, false);
        if (st != BH_OK) return st;
    }
    const float hdt = 0.5f * dt;
    for (int k = 0; k < steps; ++k) {
        const int cur = c->cur;
        phase_begin(c, 6);
        launch_kick_drift(c->stream, c->x[cur], c->y[cur], c->vx[cur], c->vy[cur], c->ax, c->ay,
                          c->n, hdt, dt);
        phase_end(c, 6);
        st = do_build(c, nullptr);
        if (st == BH_OK) st = do_forces(c, theta, false);
        if (st != BH_OK) return st;
        const int nc = c->cur;
        phase_begin(c, 6);
        launch_kick(c->stream, c->vx[nc], c->vy[nc], c->ax, c->ay, c->n, hdt);
        phase_end(c, 6);
    }
    st = launch_status(c, "leapfrog launch");
    if (st != BH_OK) return st;
    return finish_build_call(c, "leapfrog");
}

// one kick-drift-kick step on the context's stream (no host synchronisation, no profiling)
static bh_status leapfrog_step(bh_ctx* c, float dt, double theta) {
    const float hdt = 0.5f * dt;
    const int cur = c->cur;
    launch_kick_drift(c->stream, c->x[cur], c->y[cur], c->vx[cur], c->vy[cur], c->ax, c->ay, c->n,
                      hdt, dt);
    bh_status st = do_build(c, nullptr);
    if (st == BH_OK) st = do_forces(c, theta, false);
    if (st != BH_OK) return st;
    const int nc = c->cur;
    launch_kick(c->stream, c->vx[nc], c->vy[nc], c->ax, c->ay, c->n, hdt);
    return launch_status(c, "leapfrog step");
}

static void graph_reset(bh_ctx* c) {
    for (int p = 0; p < 2; ++p) {
        if (c->gexec[p]) (void)hipGraphExecDestroy(c->gexec[p]);
        if (c->graph[p]) (void)hipGraphDestroy(c->graph[p]);
        c->gexec[p] = nullptr;
        c->graph[p] = nullptr;
    }
    c->g_valid = false;
}

bh_status bh_leapfrog_graph(bh_ctx* c, int steps, float dt, double theta) {
    if (!c || steps < 0) return BH_ERR_ARG;
    if (c->n == 0) return fail(c, BH_ERR_STATE, "no particles");
    bh_status st = status_begin(c);
    if (st != BH_OK) return st;
    if (!c->have_acc) {
        st = do_build(c, nullptr);
        if (st == BH_OK) st = do_forces(c, theta, false);
        if (st != BH_OK) return st;
    }
    if (!c->g_valid || c->g_dt != dt || c->g_theta != theta || c->g_variant != c->variant) {
        graph_reset(c);
        // capture one step from each ping-pong parity; the host-side state the build updates
        // (current buffer, sort parity) is the same on replay, so it is restored afterwards
        const uint32_t prof = c->profile;
        c->profile = 0u;
        const int cur0 = c->cur;
        st = hip_check(c, hipStreamSynchronize(c->stream), "graph sync");
        for (int k = 0; k < 2 && st == BH_OK; ++k) {
            const int p = c->cur;
            st = hip_check(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal),
                           "begin capture");
            if (st != BH_OK) break;
            bh_status cs = leapfrog_step(c, dt, theta);
            hipGraph_t g = nullptr;
            st = hip_check(c, hipStreamEndCapture(c->stream, &g), "end capture");
            if (cs != BH_OK) st = cs;
            if (st == BH_OK) {
                c->graph[p] = g;
                st = hip_check(c, hipGraphInstantiate(&c->gexec[p], g, nullptr, nullptr, 0),
                               "graph instantiate");
            } else if (g) {
                (void)hipGraphDestroy(g);
            }
        }
        c->cur = cur0;
        c->profile = prof;
        if (st != BH_OK) {
            graph_reset(c);
            return st;
        }
        c->g_valid = true;
        c->g_dt = dt;
        c->g_theta = theta;
        c->g_variant = c->variant;
    }
    for (int k = 0; k < steps; ++k) {
        st = hip_check(c, hipGraphLaunch(c->gexec[c->cur], c->stream), "graph launch");
        if (st != BH_OK) return st;
        c->cur ^= 1;
    }
    c->built = true;
    c->have_acc = true;
    c->lfeat_ok = false;
    return finish_build_call(c, "graph steps");
}

bh_status bh_direct_fp64(bh_ctx* c, const uint32_t* sinks, uint32_t ns, double* ax, double* ay,
                         double* total) {
    if (!c || (!sinks && ns)) return BH_ERR_ARG;
    if (!c->built) return fail(c, BH_ERR_STATE, "no particles in sorted state (build first)");
    if (ns == 0) return BH_OK;
    for (uint32_t k = 0; k < ns; ++k)
        if (sinks[k] >= c->n) return fail(c, BH_ERR_ARG, "sink index %u out of range", sinks[k]);
    const uint32_t S = direct_splits(c->n, ns);
    const uint64_t need_s = 2ull * ns, need_p = (uint64_t)(3ull * S + 3ull) * ns;
    if (need_s > c->dsink_cap) {
        if (c->dsinks) (void)hipFree(c->dsinks);
        c->dsinks = nullptr;
        c->dsink_cap = 0;
        bh_status st = hip_check(c, hipMalloc((void**)&c->dsinks, need_s * 4), "hipMalloc sinks");
        if (st != BH_OK) return st;
        c->dsink_cap = need_s;
    }
    if (need_p > c->dpart_cap) {
        if (c->dpart) (void)hipFree(c->dpart);
        c->dpart = nullptr;
        c->dpart_cap = 0;
        bh_status st = hip_check(c, hipMalloc((void**)&c->dpart, need_p * 8), "hipMalloc partials");
        if (st != BH_OK) return st;
        c->dpart_cap = need_p;
    }
    hipStream_t s = c->stream;
    bh_status st = hip_check(c, hipMemcpyAsync(c->dsinks, sinks, (size_t)ns * 4, hipMemcpyHostToDevice, s),
                             "upload sinks");
    if (st != BH_OK) return st;
    double* out = c->dpart + (uint64_t)3 * S * ns;
    const bool grav = c->cfg.kernel == BH_KERNEL_GRAVITY;
    const double eps = c->cfg.softening;
    launch_direct_f64(s, grav ? kKernelGravity : kKernelTsne, c->pos4, c->id[c->cur], c->n,
                      c->dsinks, ns, grav ? (double)c->cfg.G : 1.0, eps * eps, c->where,
                      c->dsinks + ns, c->dpart, S, out, out + ns, out + 2ull * ns,
                      c->nsys ? c->sys_off : nullptr, c->nsys);
    st = launch_status(c, "direct launch");
    if (st != BH_OK) return st;
    const size_t b = (size_t)ns * 8;
    if (ax) st = hip_check(c, hipMemcpyAsync(ax, out, b, hipMemcpyDeviceToHost, s), "ax");
    if (st == BH_OK && ay) st = hip_check(c, hipMemcpyAsync(ay, out + ns, b, hipMemcpyDeviceToHost, s), "ay");
    if (st == BH_OK && total)
        st = hip_check(c, hipMemcpyAsync(total, out + 2 * (uint64_t)ns, b, hipMemcpyDeviceToHost, s), "tot");
    if (st != BH_OK) return st;
    return hip_check(c, hipStreamSynchronize(s), "direct");
}

bh_status bh_energy(bh_ctx* c, double theta, double* kinetic, double* potential,
                    double* momentum_x, double* momentum_y) {
    if (!c) return BH_ERR_ARG;
    if (c->cfg.kernel != BH_KERNEL_GRAVITY)
        return fail(c, BH_ERR_ARG, "energy and momentum need the gravity kernel");
    if (!c->built) return fail(c, BH_ERR_STATE, "no tree built for the current positions");
    if (c->solver != BH_SOLVER_BH) return fail(c, BH_ERR_STATE, "energy needs the Barnes-Hut walk");
    bh_status st = do_forces(c, theta, true);  // per-particle potential in tot
    if (st != BH_OK) return st;
    // the accelerations now belong to this theta: the next leapfrog call evaluates its own
    c->have_acc = false;
    const int cur = c->cur;
    launch_state_sums(c->stream, c->m[cur], c->vx[cur], c->vy[cur], c->tot, c->n, c->red_part,
                      c->dsum + 1);
    st = launch_status(c, "state sums");
    if (st != BH_OK) return st;
    double h[4];
    st = hip_check(c, hipMemcpyAsync(h, c->dsum + 1, sizeof(h), hipMemcpyDeviceToHost, c->stream),
                   "state copy");
    if (st == BH_OK) st = hip_check(c, hipStreamSynchronize(c->stream), "energy");
    if (st != BH_OK) return st;
    if (kinetic) *kinetic = h[0];
    if (potential) *potential = h[1];
    if (momentum_x) *momentum_x = h[2];
    if (momentum_y) *momentum_y = h[3];
    return BH_OK;
}

bh_status bh_direct_fp64_noalloc_async(bh_ctx* c, const uint32_t* d_sinks, uint32_t ns,
                                       double* d_ax, double* d_ay, double* d_total) {
    if (!c || (!d_sinks && ns)) return BH_ERR_ARG;
    if (!c->built) return fail(c, BH_ERR_STATE, "no particles in sorted state (build first)");
    if (ns == 0) return BH_OK;
    if (!direct_static_fits(c->n, ns))
        return fail(c, BH_ERR_CAPACITY,
                    "%u sinks exceed the static scratch of the allocation-free direct sum (at "
                    "most %u sinks)", ns, kDirectStaticSinks);
    const bool grav = c->cfg.kernel == BH_KERNEL_GRAVITY;
    const double eps = c->cfg.softening;
    launch_direct_f64_static(c->stream, grav ? kKernelGravity : kKernelTsne, c->pos4,
                             c->id[c->cur], c->n, d_sinks, ns, grav ? (double)c->cfg.G : 1.0,
                             eps * eps, c->where, c->static_scratch, d_ax, d_ay, d_total,
                             c->nsys ? c->sys_off : nullptr, c->nsys);
    return launch_status(c, "direct launch");
}

// ---------------------------------------------------------------------------------------------
// t-SNE
// ---------------------------------------------------------------------------------------------
bh_status bh_tsne_params_default(bh_tsne_params* p, uint32_t n, double theta) {
    if (!p) return BH_ERR_ARG;
    memset(p, 0, sizeof(*p));
    p->learning_rate = (double)n / 15.0;
    p->momentum_early = 0.5;
    p->momentum_late = 0.8;
    p->exag_early = 16.0;
    p->exag_late = 4.0;
    p->theta_min = theta;
    p->theta_max = theta;
    p->momentum_switch = 250;
    p->exag_switch = 250;
    p->schedule_iters = 999;
    p->repulsion_scale = 4.0;
    return BH_OK;
}

bh_status bh_tsne_params_gpu_reference(bh_tsne_params* p) {
    if (!p) return BH_ERR_ARG;
    memset(p, 0, sizeof(*p));
    p->learning_rate = 1000.0;  // tsneGpu.h:145
    p->momentum_early = 0.2;
    p->momentum_late = 0.2;
    p->exag_early = 1.0;        // P as loaded, no exaggeration (nBodySolverGpuBH.cu:430-441)
    p->exag_late = 1.0;
    p->theta_min = 1.0;         // the acceptance l / d < 1 (nBodySolverGpuBH.cu:283)
    p->theta_max = 1.0;
    p->momentum_switch = 0;
    p->exag_switch = 0;
    p->schedule_iters = 999;
    p->repulsion_scale = 1.0;   // derivative *= -1/Z (nBodySolverGpuBH.cu:424)
    return BH_OK;
}

// thetaFunction (tsne.h:208-233) runs inside updatePoints before the counter increments, so
// iteration t traverses with the value computed at t - 1; the constructor sets theta_max
// (tsne.h:105-108) for the first iteration.
static double tsne_theta(const bh_tsne_params* p, int32_t t) {
    if (t <= 0) return p->theta_max;
    double r = (double)(t - 1) / (double)(p->schedule_iters > 0 ? p->schedule_iters : 999);
    if (r > 1.0) r = 1.0;
    return p->theta_min + (0.5 + 0.5 * std::cos(M_PI * r)) * (p->theta_max - p->theta_min);
}

double bh_tsne_theta(const bh_tsne_params* p, int32_t iteration) {
    return p ? tsne_theta(p, iteration) : 0.0;
}

bh_status bh_tsne_set_p(bh_ctx* c, uint32_t n, const uint32_t* row_ptr, const uint32_t* col,
                        const float* val, uint32_t nnz) {
    if (!c || !row_ptr || (nnz && (!col || !val))) return BH_ERR_ARG;
    if (c->cfg.kernel != BH_KERNEL_TSNE) return fail(c, BH_ERR_STATE, "t-SNE needs BH_KERNEL_TSNE");
    if (n == 0 || n > c->cfg.n_max) return fail(c, BH_ERR_ARG, "n=%u outside (0, n_max]", n);
    if (row_ptr[n] != nnz) return fail(c, BH_ERR_ARG, "row_ptr[n]=%u differs from nnz=%u", row_ptr[n], nnz);
    for (uint32_t r = 0; r < n; ++r)
        if (row_ptr[r] > row_ptr[r + 1]) return fail(c, BH_ERR_ARG, "row_ptr decreases at %u", r);
    for (uint32_t e = 0; e < nnz; ++e)
        if (col[e] >= n) return fail(c, BH_ERR_ARG, "column %u out of range at %u", col[e], e);
    if (c->p_row) (void)hipFree(c->p_row);
    if (c->p_col) (void)hipFree(c->p_col);
    if (c->p_val) (void)hipFree(c->p_val);
    c->p_row = nullptr;
    c->p_col = nullptr;
    c->p_val = nullptr;
    bh_status st = hip_check(c, hipMalloc((void**)&c->p_row, (size_t)(n + 1) * 4), "hipMalloc P");
    if (st == BH_OK) st = hip_check(c, hipMalloc((void**)&c->p_col, (size_t)(nnz ? nnz : 1) * 4), "hipMalloc P");
    if (st == BH_OK) st = hip_check(c, hipMalloc((void**)&c->p_val, (size_t)(nnz ? nnz : 1) * 4), "hipMalloc P");
    if (st == BH_OK) st = hip_check(c, hipMemcpy(c->p_row, row_ptr, (size_t)(n + 1) * 4, hipMemcpyHostToDevice), "P");
    if (st == BH_OK && nnz) st = hip_check(c, hipMemcpy(c->p_col, col, (size_t)nnz * 4, hipMemcpyHostToDevice), "P");
    if (st == BH_OK && nnz) st = hip_check(c, hipMemcpy(c->p_val, val, (size_t)nnz * 4, hipMemcpyHostToDevice), "P");
    if (st != BH_OK) return st;
    c->p_n = n;
    c->p_nnz = nnz;
    return BH_OK;
}

bh_status bh_tsne_set_embedding(bh_ctx* c, const float* yx, const float* yy) {
    if (!c || !yx || !yy) return BH_ERR_ARG;
    if (c->cfg.kernel != BH_KERNEL_TSNE) return fail(c, BH_ERR_STATE, "t-SNE needs BH_KERNEL_TSNE");
    if (!c->p_n) return fail(c, BH_ERR_STATE, "set the P matrix first");
    const size_t b = (size_t)c->p_n * 4;
    bh_status st = BH_OK;
    for (int k = 0; k < 3 && st == BH_OK; ++k) {
        st = hip_check(c, hipMemcpy(c->ty_x[k], yx, b, hipMemcpyHostToDevice), "embedding");
        if (st == BH_OK) st = hip_check(c, hipMemcpy(c->ty_y[k], yy, b, hipMemcpyHostToDevice), "embedding");
    }
    if (st != BH_OK) return st;
    c->ty_cur = 0;
    c->ty_set = true;
    return BH_OK;
}

bh_status bh_tsne_get_embedding(bh_ctx* c, float* yx, float* yy) {
    if (!c) return BH_ERR_ARG;
    if (!c->ty_set) return fail(c, BH_ERR_STATE, "no embedding");
    const size_t b = (size_t)c->p_n * 4;
    bh_status st = BH_OK;
    if (yx) st = hip_check(c, hipMemcpyAsync(yx, c->ty_x[c->ty_cur], b, hipMemcpyDeviceToHost, c->stream), "y");
    if (st == BH_OK && yy) st = hip_check(c, hipMemcpyAsync(yy, c->ty_y[c->ty_cur], b, hipMemcpyDeviceToHost, c->stream), "y");
    if (st != BH_OK) return st;
    return hip_check(c, hipStreamSynchronize(c->stream), "get embedding");
}

// stage the current embedding as the solver's particles (and its interleaved copy ty_xy)
static void tsne_stage(bh_ctx* c) {
    const uint32_t n = c->p_n;
    c->n = n;
    c->cur = 0;
    launch_tsne_stage(c->stream, c->ty_x[c->ty_cur], c->ty_y[c->ty_cur], n, c->x[0], c->y[0],
                      c->m[0], c->id[0], c->ty_xy);
}

// stage the current embedding and build the tree
static bh_status tsne_stage_build(bh_ctx* c) {
    tsne_stage(c);
    return do_build(c, nullptr);
}

// The attraction needs only the staged embedding: it runs on att_stream from the stage on, while
// the tree is built and walked on the context's stream; the update waits for both.
static bh_status tsne_iteration(bh_ctx* c, const bh_tsne_params* p, int32_t t, bool keep_deriv) {
    hipStream_t s = c->stream;
    const uint32_t n = c->p_n;
    const float exag = (float)(t < p->exag_switch ? p->exag_early : p->exag_late);
    const float mu = (float)(t < p->momentum_switch ? p->momentum_early : p->momentum_late);
    tsne_stage(c);
    bh_status st = hip_check(c, hipEventRecord(c->ev_staged, s), "event");
    if (st == BH_OK) st = hip_check(c, hipStreamWaitEvent(c->att_stream, c->ev_staged, 0), "wait");
    if (st != BH_OK) return st;
    launch_tsne_attract(c->att_stream, c->p_row, c->p_col, c->p_val, n, c->ty_xy, exag, c->att_x,
                        c->att_y);
    st = hip_check(c, hipEventRecord(c->ev_att, c->att_stream), "event");
    if (st != BH_OK) return st;
    st = do_build(c, nullptr);
    if (st != BH_OK) return st;
    st = do_forces(c, tsne_theta(p, t), true);  // repulsion + per-point Z terms, Z -> dsum
    if (st != BH_OK) return st;
    launch_unsort2(s, c->id[c->cur], n, c->ax, c->ay, c->rep_x, c->rep_y);
    st = hip_check(c, hipStreamWaitEvent(s, c->ev_att, 0), "wait");
    if (st != BH_OK) return st;
    const int cur = c->ty_cur, prev = (cur + 2) % 3, next = (cur + 1) % 3;
    launch_tsne_combine(s, n, c->ty_x[cur], c->ty_y[cur], c->ty_x[prev], c->ty_y[prev], c->rep_x,
                        c->rep_y, c->att_x, c->att_y, c->dsum,
                        p->repulsion_scale > 0.0 ? p->repulsion_scale : 4.0,
                        (float)p->learning_rate, mu, c->ty_x[next], c->ty_y[next],
                        keep_deriv ? c->tder_x : nullptr, keep_deriv ? c->tder_y : nullptr);
    // rotate: next becomes current, current becomes previous
    c->ty_cur = next;
    return launch_status(c, "t-SNE step");
}

bh_status bh_tsne_step(bh_ctx* c, const bh_tsne_params* p, int32_t iteration, float* dx, float* dy,
                       double* z) {
    if (!c || !p || iteration < 0) return BH_ERR_ARG;
    if (!c->ty_set) return fail(c, BH_ERR_STATE, "set P and the embedding first");
    bh_status st = status_begin(c);
    if (st == BH_OK) st = tsne_iteration(c, p, iteration, dx || dy);
    if (st != BH_OK) return st;
    hipStream_t s = c->stream;
    const size_t b = (size_t)c->p_n * 4;
    if (dx) st = hip_check(c, hipMemcpyAsync(dx, c->tder_x, b, hipMemcpyDeviceToHost, s), "dx");
    if (st == BH_OK && dy) st = hip_check(c, hipMemcpyAsync(dy, c->tder_y, b, hipMemcpyDeviceToHost, s), "dy");
    if (st == BH_OK && z) st = hip_check(c, hipMemcpyAsync(c->h_dsum, c->dsum, 8, hipMemcpyDeviceToHost, s), "z");
    if (st != BH_OK) return st;
    st = finish_build_call(c, "t-SNE step");
    if (st == BH_OK && z) *z = *c->h_dsum;
    return st;
}

bh_status bh_tsne_run(bh_ctx* c, const bh_tsne_params* p, int32_t first, int32_t count) {
    if (!c || !p || first < 0 || count < 0) return BH_ERR_ARG;
    if (!c->ty_set) return fail(c, BH_ERR_STATE, "set P and the embedding first");
    bh_status st = status_begin(c);
    if (st != BH_OK) return st;
    for (int32_t t = first; t < first + count; ++t) {
        st = tsne_iteration(c, p, t, false);
        if (st != BH_OK) return st;
    }
    return finish_build_call(c, "t-SNE run");
}

bh_status bh_tsne_affinities(bh_ctx* c, const float* X, uint32_t n, uint32_t dim,
                             double perplexity, uint32_t k, uint32_t* row_ptr, uint32_t* col,
                             float* val, uint32_t* nnz, uint32_t* knn_idx, double* knn_d2,
                             double* beta) {
    if (!c || !X || !row_ptr || !col || !val || !nnz) return BH_ERR_ARG;
    if (n < 2 || k < 1 || k >= n || k > 128)
        return fail(c, BH_ERR_ARG, "need 1 <= k <= 128 and k < n (k = %u, n = %u)", k, n);
    if (dim < 1 || dim > 1024) return fail(c, BH_ERR_ARG, "dimension %u outside [1, 1024]", dim);
    if (!(perplexity > 0.0)) return fail(c, BH_ERR_ARG, "perplexity must be positive");
    const uint64_t nk = (uint64_t)n * k, m64 = 2 * nk;
    if (m64 >= (1ull << 31) || (uint64_t)n * n >= (1ull << 62))
        return fail(c, BH_ERR_ARG, "%u points x %u neighbours is too large", n, k);
    const uint32_t m = (uint32_t)m64;
    // scale by a power of two so that the largest |x| lands in [2^13, 2^14): no FP16 overflow.
    // The magnitude bits of a float order its |x|; infinities and NaN have all exponent bits set
    uint32_t maxbits = 0;
    {
        const uint32_t* xb = reinterpret_cast<const uint32_t*>(X);
        const uint64_t cnt = (uint64_t)n * dim;
        for (uint64_t e = 0; e < cnt; ++e) maxbits = std::max(maxbits, xb[e] & 0x7FFFFFFFu);
    }
    if (maxbits >= 0x7F800000u) return fail(c, BH_ERR_ARG, "an input coordinate is not finite");
    float maxf;
    std::memcpy(&maxf, &maxbits, 4);
    const double maxabs = maxf;
    int ex = 0;
    if (maxabs > 0.0) (void)std::frexp(maxabs, &ex);
    const float scale = maxabs > 0.0 ? std::ldexp(1.0f, 14 - ex) : 1.0f;
    const uint32_t dp = aff_dim_padded(dim), cap = aff_cap();
    hipStream_t s = c->stream;
    std::vector<void*> tmp;
    auto alloc = [&](void** p, size_t bytes) -> bh_status {
        *p = nullptr;
        hipError_t e = hipMalloc(p, bytes ? bytes : 16);
        if (e != hipSuccess) return hip_check(c, e, "hipMalloc (affinities)");
        tmp.push_back(*p);
        return BH_OK;
    };
    float *dX = nullptr, *nh = nullptr, *ah = nullptr, *eb = nullptr, *clo = nullptr, *chi = nullptr;
    float* dval = nullptr;
    _Float16* Xh = nullptr;
    uint32_t *cj = nullptr, *ccount = nullptr, *flags = nullptr, *idx = nullptr, *vals_a = nullptr,
             *vals_b = nullptr, *head = nullptr, *pos = nullptr, *spart = nullptr,
             *scratch = nullptr, *dcol = nullptr, *drp = nullptr;
    uint64_t *keys_a = nullptr, *keys_b = nullptr;
    double *d2 = nullptr, *P = nullptr, *dbeta = nullptr;
    const size_t nsp = scan_partials_size(m);
    bh_status st = BH_OK;
    st = alloc((void**)&dX, (size_t)n * dim * 4);
    if (st == BH_OK) st = alloc((void**)&Xh, (size_t)n * dp * 2);
    if (st == BH_OK) st = alloc((void**)&nh, (size_t)n * 4);
    if (st == BH_OK) st = alloc((void**)&ah, (size_t)n * 4);
    if (st == BH_OK) st = alloc((void**)&eb, (size_t)n * 4);
    if (st == BH_OK) st = alloc((void**)&cj, (size_t)n * cap * 4);
    if (st == BH_OK) st = alloc((void**)&clo, (size_t)n * cap * 4);
    if (st == BH_OK) st = alloc((void**)&chi, (size_t)n * cap * 4);
    if (st == BH_OK) st = alloc((void**)&ccount, (size_t)n * 4);
    if (st == BH_OK) st = alloc((void**)&flags, 4 * 4);  // [0] rows over capacity, [1] sort error, [2] nnz
    if (st == BH_OK) st = alloc((void**)&idx, nk * 4);
    if (st == BH_OK) st = alloc((void**)&d2, nk * 8);
    if (st == BH_OK) st = alloc((void**)&P, nk * 8);
    if (st == BH_OK) st = alloc((void**)&dbeta, (size_t)n * 8);
    if (st == BH_OK) st = alloc((void**)&keys_a, (size_t)m * 8);
    if (st == BH_OK) st = alloc((void**)&keys_b, (size_t)m * 8);
    if (st == BH_OK) st = alloc((void**)&vals_a, (size_t)m * 4);
    if (st == BH_OK) st = alloc((void**)&vals_b, (size_t)m * 4);
    if (st == BH_OK) st = alloc((void**)&head, (size_t)m * 4);
    if (st == BH_OK) st = alloc((void**)&pos, (size_t)m * 4);
    if (st == BH_OK) st = alloc((void**)&spart, (nsp + 1) * 4);
    if (st == BH_OK) st = alloc((void**)&scratch, sort_scratch_words(m) * 4);
    if (st == BH_OK) st = alloc((void**)&dcol, (size_t)m * 4);
    if (st == BH_OK) st = alloc((void**)&dval, (size_t)m * 4);
    if (st == BH_OK) st = alloc((void**)&drp, ((size_t)n + 1) * 4);
    uint32_t h_flags[4] = {0, 0, 0, 0};
    if (st == BH_OK) {
        st = hip_check(c, hipMemcpyAsync(dX, X, (size_t)n * dim * 4, hipMemcpyHostToDevice, s), "X");
    }
    if (st == BH_OK) {
        launch_fill_u32(s, flags, 4, 0u);
        launch_aff_pack(s, dX, n, dim, scale, Xh, nh, ah, eb);
        launch_aff_candidates(s, Xh, nh, ah, eb, n, dim, k, cj, clo, chi, ccount, flags);
        launch_aff_refine(s, dX, n, dim, k, cj, ccount, idx, d2);
        launch_aff_bandwidth(s, d2, n, k, std::log(perplexity), 1e-5, 200, P, dbeta);
        launch_sym_entries(s, idx, n, k, keys_a, vals_a);
        int key_bits = 1;
        while (key_bits < 64 && ((uint64_t)n * n - 1) >> key_bits) ++key_bits;
        bool in_b = false;
        launch_radix_sort(s, keys_a, vals_a, keys_b, vals_b, m, key_bits, true, scratch, flags + 1,
                          kSortSpinLimit, &in_b);
        const uint64_t* ks = in_b ? keys_b : keys_a;
        const uint32_t* vs = in_b ? vals_b : vals_a;
        launch_sym_heads(s, ks, m, head);
        launch_exclusive_scan(s, head, pos, m, spart, flags + 2);
        launch_sym_reduce(s, ks, vs, pos, m, n, (uint32_t)nk, P, dcol, dval, drp, flags + 2);
        st = launch_status(c, "affinities launch");
    }
    if (st == BH_OK)
        st = hip_check(c, hipMemcpyAsync(h_flags, flags, sizeof(h_flags), hipMemcpyDeviceToHost, s), "flags");
    if (st == BH_OK) st = hip_check(c, hipStreamSynchronize(s), "affinities");
    if (st == BH_OK && h_flags[0])
        st = fail(c, BH_ERR_CAPACITY,
                  "%u points have more than %u candidates within the FP16 error bound of their "
                  "k-th neighbour (many exact or near duplicates)", h_flags[0], cap - 64);
    if (st == BH_OK && h_flags[1]) st = fail(c, BH_ERR_HIP, "the radix sort gave up waiting");
    const uint32_t nz = h_flags[2];
    if (st == BH_OK && (nz == 0 || nz > m)) st = fail(c, BH_ERR_HIP, "bad nonzero count %u", nz);
    if (st == BH_OK) {
        *nnz = nz;
        st = hip_check(c, hipMemcpy(row_ptr, drp, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost), "row_ptr");
        if (st == BH_OK) st = hip_check(c, hipMemcpy(col, dcol, (size_t)nz * 4, hipMemcpyDeviceToHost), "col");
        if (st == BH_OK) st = hip_check(c, hipMemcpy(val, dval, (size_t)nz * 4, hipMemcpyDeviceToHost), "val");
        if (st == BH_OK && knn_idx) st = hip_check(c, hipMemcpy(knn_idx, idx, nk * 4, hipMemcpyDeviceToHost), "idx");
        if (st == BH_OK && knn_d2) st = hip_check(c, hipMemcpy(knn_d2, d2, nk * 8, hipMemcpyDeviceToHost), "d2");
        if (st == BH_OK && beta) st = hip_check(c, hipMemcpy(beta, dbeta, (size_t)n * 8, hipMemcpyDeviceToHost), "beta");
    }
    (void)hipStreamSynchronize(s);
    for (void* p : tmp) (void)hipFree(p);
    return st;
}

bh_status bh_tsne_kl(bh_ctx* c, double* kl, double* z_exact) {
    if (!c) return BH_ERR_ARG;
    if (!c->ty_set) return fail(c, BH_ERR_STATE, "set P and the embedding first");
    bh_status st = status_begin(c);
    if (st == BH_OK) st = tsne_stage_build(c);
    if (st != BH_OK) return st;
    const uint32_t n = c->p_n;
    // exact Z: FP64 direct sum over all sinks (their totals), then the KL sum over P
    const uint32_t S = direct_splits(n, n);
    const uint64_t need_s = 2ull * n, need_p = (uint64_t)(3ull * S + 3ull) * n;
    if (need_s > c->dsink_cap) {
        if (c->dsinks) (void)hipFree(c->dsinks);
        c->dsinks = nullptr;
        c->dsink_cap = 0;
        st = hip_check(c, hipMalloc((void**)&c->dsinks, need_s * 4), "hipMalloc sinks");
        if (st != BH_OK) return st;
        c->dsink_cap = need_s;
    }
    if (need_p > c->dpart_cap) {
        if (c->dpart) (void)hipFree(c->dpart);
        c->dpart = nullptr;
        c->dpart_cap = 0;
        st = hip_check(c, hipMalloc((void**)&c->dpart, need_p * 8), "hipMalloc partials");
        if (st != BH_OK) return st;
        c->dpart_cap = need_p;
    }
    hipStream_t s = c->stream;
    st = hip_check(c, hipMemcpyAsync(c->dsinks, c->iota_dev, (size_t)n * 4, hipMemcpyDeviceToDevice, s),
                   "sinks");
    if (st != BH_OK) return st;
    double* out = c->dpart + (uint64_t)3 * S * n;
    launch_direct_f64(s, kKernelTsne, c->pos4, c->id[c->cur], n, c->dsinks, n, 1.0, 0.0, c->where,
                      c->dsinks + n, c->dpart, S, out, out + n, out + 2ull * n);
    launch_sum_f64(s, out + 2ull * n, n, c->tz);
    launch_tsne_kl(s, c->p_row, c->p_col, c->p_val, n, c->ty_x[c->ty_cur], c->ty_y[c->ty_cur],
                   c->tz, c->kl_part, c->kl_out);
    st = launch_status(c, "KL");
    double h[2] = {0.0, 0.0};
    if (st == BH_OK) st = hip_check(c, hipMemcpyAsync(&h[0], c->kl_out, 8, hipMemcpyDeviceToHost, s), "kl");
    if (st == BH_OK) st = hip_check(c, hipMemcpyAsync(&h[1], c->tz, 8, hipMemcpyDeviceToHost, s), "z");
    if (st == BH_OK) st = finish_build_call(c, "KL");
    if (st != BH_OK) return st;
    if (kl) *kl = h[0];
    if (z_exact) *z_exact = h[1];
    return BH_OK;
}

bh_status bh_set_sort_spin_limit(bh_ctx* c, uint32_t polls) {
    if (!c) return BH_ERR_ARG;
    c->sort_spin = polls;
    return BH_OK;
}

bh_status bh_get_bounds(bh_ctx* c, double* bounds, double* root_side) {
    if (!c) return BH_ERR_ARG;
    if (!c->built) return fail(c, BH_ERR_STATE, "no tree built");
    bh_status st = hip_check(c, hipMemcpy(c->h_prm, c->prm, sizeof(BuildParams), hipMemcpyDeviceToHost),
                             "params");
    if (st != BH_OK) return st;
    auto d = [](uint64_t b) {
        double v;
        memcpy(&v, &b, 8);
        return v;
    };
    if (bounds) {
        bounds[0] = d(c->h_prm->minx64);
        bounds[1] = d(c->h_prm->miny64);
        bounds[2] = c->h_prm->maxx;
        bounds[3] = c->h_prm->maxy;
    }
    if (root_side) *root_side = d(c->h_prm->side64);
    return BH_OK;
}

bh_status bh_get_quantizer(bh_ctx* c, uint64_t* M, int32_t* es, uint64_t* R) {
    if (!c) return BH_ERR_ARG;
    if (!c->built) return fail(c, BH_ERR_STATE, "no tree built");
    bh_status st = hip_check(c, hipMemcpy(c->h_prm, c->prm, sizeof(BuildParams), hipMemcpyDeviceToHost),
                             "params");
    if (st != BH_OK) return st;
    if (M) *M = c->h_prm->quant.M;
    if (es) *es = c->h_prm->quant.es;
    if (R) *R = c->h_prm->quant.R;
    return BH_OK;
}

bh_status bh_get_morton(bh_ctx* c, uint64_t* codes_sorted, uint32_t* perm) {
    if (!c) return BH_ERR_ARG;
    if (!c->built) return fail(c, BH_ERR_STATE, "no tree built");
    const uint32_t n = c->n;
    bh_status st = BH_OK;
    const void* keys = c->keys_in_b ? c->keys_b : c->keys_a;
    const uint32_t* vals = c->keys_in_b ? c->vals_b : c->vals_a;
    if (codes_sorted) {
        if (c->wide) {
            st = hip_check(c, hipMemcpy(codes_sorted, keys, (size_t)n * 8, hipMemcpyDeviceToHost), "codes");
        } else {
            std::vector<uint32_t> tmp(n);
            st = hip_check(c, hipMemcpy(tmp.data(), keys, (size_t)n * 4, hipMemcpyDeviceToHost), "codes");
            for (uint32_t i = 0; i < n; ++i) codes_sorted[i] = tmp[i];
        }
    }
    if (st == BH_OK && perm) {
        // ids of the sorted state are the caller's original indices
        st = hip_check(c, hipMemcpy(perm, c->id[c->cur], (size_t)n * 4, hipMemcpyDeviceToHost), "perm");
    }
    (void)vals;
    return st;
}

bh_status bh_tree_info(bh_ctx* c, uint32_t* n_nodes, uint32_t* max_level, uint32_t* overfull) {
    if (!c) return BH_ERR_ARG;
    if (!c->built) return fail(c, BH_ERR_STATE, "no tree built");
    const int limit = c->wide ? c->B : c->depth;
    launch_tree_stats(c->stream, c->nodes, c->prm, c->cap_nodes, limit,
                      (uint32_t)c->cfg.max_children, c->stats_part, kStatsBlocks);
    bh_status st = launch_status(c, "tree stats");
    if (st == BH_OK)
        st = hip_check(c, hipMemcpyAsync(c->h_stats, c->stats_part, kStatsBlocks * sizeof(uint2),
                                         hipMemcpyDeviceToHost, c->stream), "stats");
    if (st == BH_OK)
        st = hip_check(c, hipMemcpyAsync(c->h_prm, c->prm, sizeof(BuildParams), hipMemcpyDeviceToHost,
                                         c->stream), "params");
    if (st == BH_OK) st = hip_check(c, hipStreamSynchronize(c->stream), "tree info");
    if (st != BH_OK) return st;
    uint32_t mx = 0, of = 0;
    for (uint32_t b = 0; b < kStatsBlocks; ++b) {
        mx = std::max(mx, c->h_stats[b].x);
        of += c->h_stats[b].y;
    }
    if (n_nodes) *n_nodes = c->h_prm->n_nodes;
    if (max_level) *max_level = mx;
    if (overfull) *overfull = of;
    return BH_OK;
}

bh_status bh_get_tree(bh_ctx* c, int32_t* level, uint32_t* first, uint32_t* count, uint32_t* skip,
                      uint8_t* flags, float* cx, float* cy, float* mass) {
    uint32_t nn = 0;
    bh_status st = bh_tree_info(c, &nn, nullptr, nullptr);
    if (st == BH_OK && !c->moments) {
        // the node records' moments are made on the context's stream, which the copy below
        // (on the null stream) does not wait for
        st = ensure_moments(c);
        if (st == BH_OK) st = hip_check(c, hipStreamSynchronize(c->stream), "upward");
    }
    if (st != BH_OK) return st;
    std::vector<Node> h(nn);
    if (nn) {
        st = hip_check(c, hipMemcpy(h.data(), c->nodes, (size_t)nn * sizeof(Node), hipMemcpyDeviceToHost),
                       "nodes");
        if (st != BH_OK) return st;
    }
    for (uint32_t k = 0; k < nn; ++k) {
        if (level) level[k] = (int32_t)(h[k].meta & kMetaLevelMask);
        if (first) first[k] = h[k].first;
        if (count) count[k] = h[k].count;
        if (skip) skip[k] = h[k].skip;
        if (flags) flags[k] = (uint8_t)(((h[k].meta & kMetaLeaf) ? 1 : 0) | ((h[k].meta & kMetaPP) ? 2 : 0));
        if (cx) cx[k] = h[k].cx;
        if (cy) cy[k] = h[k].cy;
        if (mass) mass[k] = h[k].mass;
    }
    return BH_OK;
}

bh_status bh_get_tree_slots(bh_ctx* c, int32_t* depth, uint64_t* n_slots, uint32_t* first_child,
                            uint32_t* first_particle, uint32_t* count) {
    if (!c) return BH_ERR_ARG;
    if (c->wide) return fail(c, BH_ERR_STATE, "slot layout exists only in MORTON_FIXED mode");
    if (!c->built) return fail(c, BH_ERR_STATE, "no tree built");
    const uint64_t slots = ((1ull << (2 * (c->depth + 1))) - 1) / 3;
    if (depth) *depth = c->depth;
    if (n_slots) *n_slots = slots;
    if (!first_child && !first_particle && !count) return BH_OK;
    const void* keys = c->keys_in_b ? c->keys_b : c->keys_a;
    launch_tree_slots(c->stream, (const uint32_t*)keys, c->nodes, c->prm, c->cap_nodes, c->depth,
                      c->slot_a, c->slot_b, c->slot_c, slots);
    bh_status st = launch_status(c, "slots");
    if (st != BH_OK) return st;
    const size_t b = slots * 4;
    if (first_child) st = hip_check(c, hipMemcpyAsync(first_child, c->slot_a, b, hipMemcpyDeviceToHost, c->stream), "slots");
    if (st == BH_OK && first_particle)
        st = hip_check(c, hipMemcpyAsync(first_particle, c->slot_b, b, hipMemcpyDeviceToHost, c->stream), "slots");
    if (st == BH_OK && count)
        st = hip_check(c, hipMemcpyAsync(count, c->slot_c, b, hipMemcpyDeviceToHost, c->stream), "slots");
    if (st != BH_OK) return st;
    return hip_check(c, hipStreamSynchronize(c->stream), "slots");
}

bh_status bh_counters(bh_ctx* c, uint64_t* pn, uint64_t* pp) {
    if (!c) return BH_ERR_ARG;
    if (!c->have_acc) return fail(c, BH_ERR_STATE, "no force evaluation yet");
    const uint32_t nb = c->trav_blocks * kWavesPerBlock;  // waves of the last walk
    bh_status st = hip_check(c, hipMemcpyAsync(c->h_counters, c->counters,
                                               2ull * nb * sizeof(unsigned long long),
                                               hipMemcpyDeviceToHost, c->stream), "counters");
    if (st == BH_OK) st = hip_check(c, hipStreamSynchronize(c->stream), "counters");
    if (st != BH_OK) return st;
    uint64_t a = 0, b = 0;
    for (uint32_t k = 0; k < nb; ++k) {
        a += c->h_counters[2 * k];
        b += c->h_counters[2 * k + 1];
    }
    if (pn) *pn = a;
    if (pp) *pp = b;
    return BH_OK;
}

bh_status bh_profile_enable(bh_ctx* c, int enable) {
    if (!c) return BH_ERR_ARG;
    c->profile = enable ? (1u << BH_NUM_PHASES) - 1u : 0u;
    return BH_OK;
}

bh_status bh_profile_phases(bh_ctx* c, uint32_t mask) {
    if (!c) return BH_ERR_ARG;
    if (mask >> BH_NUM_PHASES) return fail(c, BH_ERR_ARG, "phase mask 0x%x has bits beyond phase %d", mask, BH_NUM_PHASES - 1);
    c->profile = mask;
    return BH_OK;
}

bh_status bh_profile_read(bh_ctx* c, double* ms, uint64_t* launches, int reset) {
    if (!c) return BH_ERR_ARG;
    bh_status st = hip_check(c, hipStreamSynchronize(c->stream), "profile sync");
    if (st != BH_OK) return st;
    phase_collect(c);
    for (int p = 0; p < BH_NUM_PHASES; ++p) {
        if (ms) ms[p] = c->phases[p].ms;
        if (launches) launches[p] = c->phases[p].launches;
        if (reset) {
            c->phases[p].ms = 0.0;
            c->phases[p].launches = 0;
        }
    }
    return BH_OK;
}

bh_status bh_set_traverse_variant(bh_ctx* c, int variant) {
    if (!c) return BH_ERR_ARG;
    if (variant < 0 || variant >= kTraverseVariants)
        return fail(c, BH_ERR_ARG, "traversal variant %d outside [0, %d)", variant, kTraverseVariants);
    c->variant = variant;
    return BH_OK;
}

int bh_traverse_variant_count(void) { return kTraverseVariants; }

bh_status bh_set_solver(bh_ctx* c, int solver) {
    if (!c) return BH_ERR_ARG;
    if (solver != BH_SOLVER_BH && solver != BH_SOLVER_FMM_SYM)
        return fail(c, BH_ERR_ARG, "unknown solver %d", solver);
    if (solver == BH_SOLVER_FMM_SYM) {
        if (c->cfg.tree_mode != BH_TREE_MORTON_FIXED || c->nsys || c->cfg.kernel != BH_KERNEL_TSNE)
            return fail(c, BH_ERR_ARG,
                        "the FMM solver runs the t-SNE kernels on the fixed-depth Morton tree: "
                        "one system, BH_KERNEL_TSNE, BH_TREE_MORTON_FIXED");
        if (!c->fmm_buf) {
            const uint64_t bytes = fmm_workspace_bytes(c->slot_cap, c->cfg.n_max);
            bh_status st = hip_check(c, hipMalloc(&c->fmm_buf, bytes), "hipMalloc FMM buffers");
            if (st != BH_OK) return st;
        }
    }
    if (solver != c->solver) c->g_valid = false;
    c->solver = solver;
    c->have_acc = false;
    return BH_OK;
}

// Uploads work table v: `order` lists original entries (h_work0[v]) heaviest first; the first
// `nh` of them are walked as kHeavySlices slices each.
static bh_status upload_work(bh_ctx* c, int v, const std::vector<uint32_t>& order, uint32_t nh) {
    const std::vector<uint4>& w0 = c->h_work0[v];
    if (nh > c->side_cap) {  // grown in place of the old buffer: 3 planes x nh x slices x sinks
        bh_status st = hip_check(c, hipStreamSynchronize(c->stream), "side sync");
        if (st != BH_OK) return st;
        if (c->side) (void)hipFree(c->side);
        c->side = nullptr;
        c->side_cap = 0;
        st = hip_check(c, hipMalloc((void**)&c->side, 3ull * nh * kHeavySlices * kHeavySinks * 4),
                       "hipMalloc heavy slices");
        if (st != BH_OK) return st;
        c->side_cap = nh;
    }
    std::vector<uint4> out, hv(nh);
    out.reserve(w0.size() + (size_t)nh * (kHeavySlices - 1));
    for (uint32_t h = 0; h < nh; ++h) {
        const uint4 e = w0[order[h]];
        hv[h] = make_uint4(e.x, e.y, kHeavySlices, 0u);
        for (uint32_t f = 0; f < kHeavySlices; ++f)
            out.push_back(make_uint4(e.x, e.y, e.z, (h << 8) | (kHeavySlices << 4) | f));
    }
    for (size_t i = nh; i < order.size(); ++i) out.push_back(w0[order[i]]);
    bh_status st = hip_check(c, hipMemcpyAsync(c->work[v], out.data(), out.size() * sizeof(uint4),
                                               hipMemcpyHostToDevice, c->stream), "work");
    if (st == BH_OK && nh)
        st = hip_check(c, hipMemcpyAsync(c->heavy[v], hv.data(), nh * sizeof(uint4),
                                         hipMemcpyHostToDevice, c->stream), "heavy");
    if (st == BH_OK) st = hip_check(c, hipStreamSynchronize(c->stream), "work table");
    if (st != BH_OK) return st;
    c->n_work[v] = (uint32_t)out.size();
    c->nheavy[v] = nh;
    c->g_valid = false;  // captured steps hold the old grid
    return BH_OK;
}

bh_status bh_batch_order_by_cost(bh_ctx* c) {
    if (!c) return BH_ERR_ARG;
    if (!c->nsys) return fail(c, BH_ERR_STATE, "not a batched context");
    if (!c->have_acc) return fail(c, BH_ERR_STATE, "no force evaluation yet");
    if (c->split != 1) return fail(c, BH_ERR_STATE, "order by cost after a walk with split 1");
    const int v = variant_ns_index(c->variant);
    const uint32_t nw = c->n_work[v], nb = (nw + kWavesPerBlock - 1) / kWavesPerBlock;
    if (c->trav_blocks != nb)
        return fail(c, BH_ERR_STATE, "the last walk used another variant's work table");
    std::vector<unsigned long long> cnt(2ull * nw);  // per wave: entry e's pair
    std::vector<uint4> wk(nw);
    bh_status st = hip_check(c, hipMemcpyAsync(cnt.data(), c->counters, cnt.size() * 8,
                                               hipMemcpyDeviceToHost, c->stream), "counters");
    if (st == BH_OK)
        st = hip_check(c, hipMemcpyAsync(wk.data(), c->work[v], (size_t)nw * sizeof(uint4),
                                         hipMemcpyDeviceToHost, c->stream), "work");
    if (st == BH_OK) st = hip_check(c, hipStreamSynchronize(c->stream), "order by cost");
    if (st != BH_OK) return st;
    // cost of each original entry: its wave's interactions in the last walk (a heavy entry's
    // slices add up; entries are identified by their first sink), per lane: the wave's walk is
    // shared by its lanes, so its time follows a lane's sinks (sink k of an entry sits on lane
    // k mod 64), not the entry's sink count
    const std::vector<uint4>& w0 = c->h_work0[v];
    std::vector<std::pair<uint32_t, uint32_t>> by_first(w0.size());
    for (uint32_t i = 0; i < w0.size(); ++i) by_first[i] = {w0[i].x, i};
    std::sort(by_first.begin(), by_first.end());
    std::vector<unsigned long long> cost(w0.size(), 0ull);
    for (uint32_t e = 0; e < nw; ++e) {
        const auto it = std::lower_bound(by_first.begin(), by_first.end(),
                                         std::make_pair(wk[e].x, 0u));
        if (it == by_first.end() || it->first != wk[e].x) continue;
        cost[it->second] += cnt[2 * e] + cnt[2 * e + 1];
    }
    for (uint32_t i = 0; i < w0.size(); ++i)
        if (w0[i].y) cost[i] = cost[i] * ((w0[i].y + kWave - 1) / kWave) / w0[i].y;
    std::vector<std::pair<unsigned long long, uint32_t>> ord(w0.size());  // (~cost, entry)
    unsigned long long total = 0;
    for (uint32_t i = 0; i < w0.size(); ++i) {
        ord[i] = {~cost[i], i};
        total += cost[i];
    }
    std::sort(ord.begin(), ord.end());  // descending cost, ties by entry
    std::vector<uint32_t> order(w0.size());
    for (uint32_t i = 0; i < w0.size(); ++i) order[i] = ord[i].second;
    // heavy: the costliest entries, at least kHeavyFactor times the mean, at most kHeavyMax of
    // them (the 256-system workload: a few thousand entries whose sinks sit under one-child
    // chains of thousands of occupants, 15-50 times the mean cost, 4-17 ms each of a 17 ms walk
    // on too few waves to fill the device). Each of their slices repeats the path's acceptance tests
    // but takes an eighth of the monopoles and of the particle-particle sums.
    constexpr unsigned long long kHeavyFactor = 16;
    uint32_t nh = 0;
    while (nh < std::min<size_t>(kHeavyMax, order.size()) && cost[order[nh]] > 0 &&
           cost[order[nh]] * w0.size() >= kHeavyFactor * total)
        ++nh;
    return upload_work(c, v, order, nh);
}

bh_status bh_set_walk_priority(bh_ctx* c, int on) {
    if (!c) return BH_ERR_ARG;
    if ((on != 0) != c->walk_prio) c->g_valid = false;  // captured steps hold the old walk
    c->walk_prio = on != 0;
    return BH_OK;
}

bh_status bh_get_walk_priority(bh_ctx* c, int* on) {
    if (!c || !on) return BH_ERR_ARG;
    *on = c->walk_prio ? 1 : 0;
    return BH_OK;
}

bh_status bh_set_traverse_split(bh_ctx* c, int split) {
    if (!c) return BH_ERR_ARG;
    if (split < 1 || split > (int)kMaxSplit)
        return fail(c, BH_ERR_ARG, "traversal split %d outside [1, %u]", split, kMaxSplit);
    if (split > 1 && (uint32_t)split > c->part_split) {
        // grown in place of the old buffer: 3 outputs x F parts x n_max
        bh_status st = hip_check(c, hipStreamSynchronize(c->stream), "split sync");
        if (st != BH_OK) return st;
        if (c->part) (void)hipFree(c->part);
        c->part = nullptr;
        c->part_split = 0;
        st = hip_check(c, hipMalloc(&c->part, 3ull * split * c->cfg.n_max * sizeof(float)),
                       "hipMalloc split parts");
        if (st != BH_OK) {
            c->split = 1;
            return st;
        }
        c->part_split = (uint32_t)split;
    }
    if (split > 1 && c->nsys) {
        // heavy-entry slices and split walks do not combine: back to the tables as built
        for (int v = 0; v < 3; ++v) {
            if (!c->nheavy[v]) continue;
            std::vector<uint32_t> order(c->h_work0[v].size());
            for (uint32_t i = 0; i < order.size(); ++i) order[i] = i;
            bh_status st = upload_work(c, v, order, 0);
            if (st != BH_OK) return st;
        }
    }
    if ((uint32_t)split != c->split) c->g_valid = false;  // captured steps hold the old walk
    c->split = (uint32_t)split;
    return BH_OK;
}

bh_status bh_synchronize(bh_ctx* c) {
    if (!c) return BH_ERR_ARG;
    return hip_check(c, hipStreamSynchronize(c->stream), "synchronize");
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// learned far-field correction
// ---------------------------------------------------------------------------------------------
struct bh_mlp {
    std::string err;
    hipStream_t stream = nullptr;
    float* params = nullptr;  // FP32 master copy
    float *m1 = nullptr, *m2 = nullptr;
    MlpBf16* wbf = nullptr;
    int step = 0;
    int terms = kLearnTerms;  // basis terms of the loss and of the walk's correction (3 or 7)
    float *X = nullptr, *T = nullptr;  // data set
    uint32_t cap = 0, n = 0;
    float* S = nullptr;  // training scratch
    size_t S_cap = 0;
    float* partials = nullptr;
    size_t part_cap = 0;
    double* loss_part = nullptr;
    size_t loss_cap = 0;
    uint32_t* order = nullptr;
    std::vector<void*> allocs;
};

namespace {

bh_status mfail(bh_mlp* m, bh_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (m) m->err = buf;
    return st;
}

bh_status mcheck(bh_mlp* m, hipError_t e, const char* what) {
    if (e == hipSuccess) return BH_OK;
    return mfail(m, e == hipErrorOutOfMemory ? BH_ERR_OOM : BH_ERR_HIP, "%s: %s", what,
                 hipGetErrorString(e));
}

template <typename T>
bh_status malloc_m(bh_mlp* m, T** p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) ? count * sizeof(T) : 16);
    if (e != hipSuccess) return mcheck(m, e, "hipMalloc");
    m->allocs.push_back(q);
    *p = (T*)q;
    return BH_OK;
}

// grow-only scratch owned by the network object (freed with it)
template <typename T>
bh_status ensure_m(bh_mlp* m, T** p, size_t* cap, size_t count) {
    if (count <= *cap) return BH_OK;
    if (*p) {
        for (auto& a : m->allocs)
            if (a == (void*)*p) a = nullptr;
        (void)hipFree(*p);
        *p = nullptr;
        *cap = 0;
    }
    bh_status st = malloc_m(m, p, count);
    if (st == BH_OK) *cap = count;
    return st;
}

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

bh_status mlp_upload_params(bh_mlp* m, const float* host) {
    bh_status st = mcheck(m, hipMemcpyAsync(m->params, host, kMlpParams * 4, hipMemcpyHostToDevice,
                                            m->stream), "params");
    if (st != BH_OK) return st;
    st = mcheck(m, hipMemsetAsync(m->m1, 0, kMlpParams * 4, m->stream), "adam");
    if (st == BH_OK) st = mcheck(m, hipMemsetAsync(m->m2, 0, kMlpParams * 4, m->stream), "adam");
    m->step = 0;
    launch_mlp_prep(m->stream, m->params, m->wbf);
    if (st == BH_OK) st = mcheck(m, hipGetLastError(), "prep");
    if (st == BH_OK) st = mcheck(m, hipStreamSynchronize(m->stream), "params");
    return st;
}

// host buffer -> temporary device buffer (caller frees)
template <typename T>
bh_status to_device(bh_mlp* m, const T* h, size_t count, T** d) {
    *d = nullptr;
    hipError_t e = hipMalloc((void**)d, count * sizeof(T) ? count * sizeof(T) : 16);
    if (e != hipSuccess) return mcheck(m, e, "hipMalloc");
    if (count) e = hipMemcpyAsync(*d, h, count * sizeof(T), hipMemcpyHostToDevice, m->stream);
    return mcheck(m, e, "upload");
}

}  // namespace

extern "C" {

uint32_t bh_mlp_param_count(void) { return kMlpParams; }

const char* bh_mlp_last_error(const bh_mlp* m) { return m ? m->err.c_str() : "null network"; }

void bh_mlp_destroy(bh_mlp* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void* p : m->allocs)
        if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

bh_status bh_mlp_create(bh_mlp** out, uint32_t max_samples, uint64_t seed) {
    if (!out) return BH_ERR_ARG;
    *out = nullptr;
    bh_mlp* m = new (std::nothrow) bh_mlp();
    if (!m) return BH_ERR_OOM;
    bh_status st = mcheck(m, hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking), "stream");
    if (st == BH_OK) st = malloc_m(m, &m->params, kMlpParams);
    if (st == BH_OK) st = malloc_m(m, &m->m1, kMlpParams);
    if (st == BH_OK) st = malloc_m(m, &m->m2, kMlpParams);
    if (st == BH_OK) st = malloc_m(m, &m->wbf, 1);
    if (st == BH_OK) st = malloc_m(m, &m->X, (size_t)max_samples * kLearnSample);
    if (st == BH_OK) st = malloc_m(m, &m->T, (size_t)max_samples * 2);
    if (st == BH_OK) st = malloc_m(m, &m->order, max_samples);
    if (st != BH_OK) {
        *out = m;  // error string readable; caller destroys
        return st;
    }
    m->cap = max_samples;
    // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases
    std::vector<float> h(kMlpParams, 0.f);
    uint64_t s = seed;
    auto fill = [&](int off, int fan_in, int count) {
        const double a = 1.0 / std::sqrt((double)fan_in);
        for (int i = 0; i < count; ++i) {
            const double u = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
            h[off + i] = (float)((2.0 * u - 1.0) * a);
        }
    };
    fill(kMlpOffW1, kMlpIn, kMlpIn * kMlpH);
    fill(kMlpOffW2, kMlpH, kMlpH * kMlpH);
    fill(kMlpOffW3, kMlpH, kMlpH * kMlpOut);
    st = mlp_upload_params(m, h.data());
    *out = m;
    return st;
}

bh_status bh_mlp_set_params(bh_mlp* m, const float* params, uint32_t count) {
    if (!m || !params) return BH_ERR_ARG;
    if (count != (uint32_t)kMlpParams) return mfail(m, BH_ERR_ARG, "expected %d parameters", kMlpParams);
    return mlp_upload_params(m, params);
}

bh_status bh_mlp_set_terms(bh_mlp* m, int32_t terms) {
    if (!m) return BH_ERR_ARG;
    if (terms != 3 && terms != kLearnTerms)
        return mfail(m, BH_ERR_ARG, "terms must be 3 or %d", kLearnTerms);
    m->terms = terms;
    return BH_OK;
}

int32_t bh_mlp_get_terms(const bh_mlp* m) { return m ? m->terms : 0; }

bh_status bh_mlp_get_params(bh_mlp* m, float* params, uint32_t count) {
    if (!m || !params) return BH_ERR_ARG;
    if (count != (uint32_t)kMlpParams) return mfail(m, BH_ERR_ARG, "expected %d parameters", kMlpParams);
    bh_status st = mcheck(m, hipMemcpyAsync(params, m->params, kMlpParams * 4, hipMemcpyDeviceToHost,
                                            m->stream), "params");
    if (st == BH_OK) st = mcheck(m, hipStreamSynchronize(m->stream), "params");
    return st;
}

bh_status bh_mlp_forward(bh_mlp* m, const float* X, uint32_t n, float* Y, int bf16_mfma) {
    if (!m || (n && (!X || !Y))) return BH_ERR_ARG;
    if (!n) return BH_OK;
    float *dX = nullptr, *dY = nullptr;
    bh_status st = to_device(m, X, (size_t)n * kMlpIn, &dX);
    if (st == BH_OK) st = mcheck(m, hipMalloc((void**)&dY, (size_t)n * kMlpOut * 4), "hipMalloc");
    if (st == BH_OK) {
        if (bf16_mfma) launch_mlp_infer_bf16(m->stream, dX, n, m->wbf, dY);
        else launch_mlp_forward_f32(m->stream, dX, n, m->params, dY);
        st = mcheck(m, hipGetLastError(), "forward");
    }
    if (st == BH_OK)
        st = mcheck(m, hipMemcpyAsync(Y, dY, (size_t)n * kMlpOut * 4, hipMemcpyDeviceToHost, m->stream), "Y");
    if (st == BH_OK) st = mcheck(m, hipStreamSynchronize(m->stream), "forward");
    (void)hipStreamSynchronize(m->stream);
    if (dX) (void)hipFree(dX);
    if (dY) (void)hipFree(dY);
    return st;
}

static bh_status mlp_grad_device(bh_mlp* m, const float* dX, const float* dT, const uint32_t* order,
                                 uint32_t B) {
    bh_status st = ensure_m(m, &m->S, &m->S_cap, mlp_scratch_floats(B));
    if (st == BH_OK) st = ensure_m(m, &m->partials, &m->part_cap, (size_t)mlp_grad_chunks(B) * kMlpParams);
    if (st == BH_OK) st = ensure_m(m, &m->loss_part, &m->loss_cap, mlp_loss_blocks(B));
    if (st != BH_OK) return st;
    launch_mlp_grad(m->stream, dX, dT, order, B, m->params, m->S, m->loss_part, m->partials,
                    m->terms);
    return mcheck(m, hipGetLastError(), "gradient");
}

bh_status bh_mlp_loss_grad(bh_mlp* m, const float* X, const float* T, uint32_t n, float* grads,
                           double* loss) {
    if (!m || !X || !T || !n) return BH_ERR_ARG;
    float *dX = nullptr, *dT = nullptr, *dG = nullptr;
    bh_status st = to_device(m, X, (size_t)n * kLearnSample, &dX);
    if (st == BH_OK) st = to_device(m, T, (size_t)n * 2, &dT);
    if (st == BH_OK) st = mcheck(m, hipMalloc((void**)&dG, kMlpParams * 4), "hipMalloc");
    if (st == BH_OK) st = mlp_grad_device(m, dX, dT, nullptr, n);
    if (st == BH_OK) {
        launch_mlp_adam(m->stream, m->params, m->partials, mlp_grad_chunks(n), m->m1, m->m2, 0.f, 1,
                        dG, 0);
        st = mcheck(m, hipGetLastError(), "gradient sum");
    }
    std::vector<double> lp(mlp_loss_blocks(n));
    if (st == BH_OK && grads)
        st = mcheck(m, hipMemcpyAsync(grads, dG, kMlpParams * 4, hipMemcpyDeviceToHost, m->stream), "grads");
    if (st == BH_OK)
        st = mcheck(m, hipMemcpyAsync(lp.data(), m->loss_part, lp.size() * 8, hipMemcpyDeviceToHost,
                                      m->stream), "loss");
    if (st == BH_OK) st = mcheck(m, hipStreamSynchronize(m->stream), "gradient");
    (void)hipStreamSynchronize(m->stream);
    if (dX) (void)hipFree(dX);
    if (dT) (void)hipFree(dT);
    if (dG) (void)hipFree(dG);
    if (st == BH_OK && loss) {
        double s = 0.0;
        for (double v : lp) s += v;
        *loss = s / (2.0 * n);  // mean over the n * 2 outputs
    }
    return st;
}

bh_status bh_mlp_adam_step(bh_mlp* m, const float* grads, float lr) {
    if (!m || !grads) return BH_ERR_ARG;
    float* dG = nullptr;
    bh_status st = to_device(m, grads, kMlpParams, &dG);
    if (st == BH_OK) {
        launch_mlp_adam_grad(m->stream, m->params, dG, m->m1, m->m2, lr, ++m->step);
        launch_mlp_prep(m->stream, m->params, m->wbf);
        st = mcheck(m, hipGetLastError(), "adam");
    }
    if (st == BH_OK) st = mcheck(m, hipStreamSynchronize(m->stream), "adam");
    (void)hipStreamSynchronize(m->stream);
    if (dG) (void)hipFree(dG);
    return st;
}

bh_status bh_mlp_dataset_clear(bh_mlp* m) {
    if (!m) return BH_ERR_ARG;
    m->n = 0;
    return BH_OK;
}

bh_status bh_mlp_dataset_size(bh_mlp* m, uint32_t* n) {
    if (!m || !n) return BH_ERR_ARG;
    *n = m->n;
    return BH_OK;
}

bh_status bh_mlp_dataset_get(bh_mlp* m, float* X, float* T, uint32_t n) {
    if (!m) return BH_ERR_ARG;
    if (n > m->n) return mfail(m, BH_ERR_ARG, "data set holds %u samples", m->n);
    bh_status st = BH_OK;
    if (X) st = mcheck(m, hipMemcpyAsync(X, m->X, (size_t)n * kLearnSample * 4, hipMemcpyDeviceToHost, m->stream), "X");
    if (st == BH_OK && T)
        st = mcheck(m, hipMemcpyAsync(T, m->T, (size_t)n * 2 * 4, hipMemcpyDeviceToHost, m->stream), "T");
    if (st == BH_OK) st = mcheck(m, hipStreamSynchronize(m->stream), "data set");
    return st;
}

bh_status bh_mlp_dataset_set(bh_mlp* m, const float* X, const float* T, uint32_t n) {
    if (!m || (n && (!X || !T))) return BH_ERR_ARG;
    if (n > m->cap) return mfail(m, BH_ERR_CAPACITY, "%u samples exceed the capacity %u", n, m->cap);
    bh_status st = BH_OK;
    if (n) st = mcheck(m, hipMemcpyAsync(m->X, X, (size_t)n * kLearnSample * 4, hipMemcpyHostToDevice, m->stream), "X");
    if (st == BH_OK && n)
        st = mcheck(m, hipMemcpyAsync(m->T, T, (size_t)n * 2 * 4, hipMemcpyHostToDevice, m->stream), "T");
    if (st == BH_OK) st = mcheck(m, hipStreamSynchronize(m->stream), "data set");
    if (st == BH_OK) m->n = n;
    return st;
}

bh_status bh_mlp_train(bh_mlp* m, int epochs, uint32_t batch, float lr, uint64_t seed,
                       double* final_loss) {
    if (!m || epochs < 0 || batch == 0) return BH_ERR_ARG;
    if (m->n == 0) return mfail(m, BH_ERR_STATE, "empty data set");
    const uint32_t n = m->n;
    const uint32_t B = std::min(batch, n);
    std::vector<uint32_t> perm(n);
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    uint64_t s = seed;
    bh_status st = BH_OK;
    for (int ep = 0; ep < epochs && st == BH_OK; ++ep) {
        for (uint32_t i = n - 1; i > 0; --i) {  // Fisher-Yates
            const uint32_t j = (uint32_t)(splitmix64(s) % (uint64_t)(i + 1));
            std::swap(perm[i], perm[j]);
        }
        st = mcheck(m, hipMemcpyAsync(m->order, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice,
                                      m->stream), "order");
        for (uint32_t b0 = 0; b0 + B <= n && st == BH_OK; b0 += B) {  // full batches only
            st = mlp_grad_device(m, m->X, m->T, m->order + b0, B);
            if (st != BH_OK) break;
            launch_mlp_adam(m->stream, m->params, m->partials, mlp_grad_chunks(B), m->m1, m->m2, lr,
                            ++m->step, nullptr, 1);
            st = mcheck(m, hipGetLastError(), "adam");
        }
        // the host permutation buffer is rewritten next epoch: wait for the upload to be used
        if (st == BH_OK) st = mcheck(m, hipStreamSynchronize(m->stream), "epoch");
    }
    if (st == BH_OK) {
        launch_mlp_prep(m->stream, m->params, m->wbf);
        st = mcheck(m, hipGetLastError(), "prep");
    }
    if (st == BH_OK && final_loss) {
        // loss over the whole data set in chunks (forward + backward kernels, loss partials only)
        double tot = 0.0;
        const uint32_t C = 65536;
        for (uint32_t o = 0; o < n && st == BH_OK; o += C) {
            const uint32_t b = std::min(C, n - o);
            st = mlp_grad_device(m, m->X + (size_t)o * kLearnSample, m->T + (size_t)o * 2, nullptr, b);
            std::vector<double> lp(mlp_loss_blocks(b));
            if (st == BH_OK)
                st = mcheck(m, hipMemcpyAsync(lp.data(), m->loss_part, lp.size() * 8,
                                              hipMemcpyDeviceToHost, m->stream), "loss");
            if (st == BH_OK) st = mcheck(m, hipStreamSynchronize(m->stream), "loss");
            for (double v : lp) tot += v;
        }
        *final_loss = tot / (2.0 * n);
    }
    if (st == BH_OK) st = mcheck(m, hipStreamSynchronize(m->stream), "train");
    return st;
}

}  // extern "C"

namespace {

// cell features of the current tree (after the upward pass: they use the second moments)
bh_status ensure_features(bh_ctx* c) {
    if (ensure_moments(c) != BH_OK) return BH_ERR_HIP;
    if (c->lfeat_ok) return BH_OK;
    const void* codes = c->keys_in_b ? c->keys_b : c->keys_a;
    phase_begin(c, 4);
    launch_cell_features(c->stream, c->pos4, c->nodes, c->m2, codes, c->wide, c->B, c->prm,
                         c->cap_nodes, c->lfeat, c->lbig, c->cfg.n_max);
    phase_end(c, 4);
    bh_status st = launch_status(c, "cell features");
    if (st == BH_OK) c->lfeat_ok = true;
    return st;
}

// the network's coefficient record of every cell, for the walk
bh_status ensure_coefs(bh_ctx* c, const bh_mlp* m) {
    bh_status st = ensure_features(c);
    if (st != BH_OK) return st;
    launch_cell_coefs(c->stream, c->lfeat, c->nodes, c->prm, c->cap_nodes, m->wbf, c->lcoef);
    return launch_status(c, "cell coefficients");
}

}  // namespace

extern "C" {

bh_status bh_learned_enable(bh_ctx* c, int enable) {
    if (!c) return BH_ERR_ARG;
    if (c->cfg.kernel != BH_KERNEL_GRAVITY)
        return fail(c, BH_ERR_STATE, "the learned correction is trained for the gravity kernel");
    if (c->cfg.tree_mode != BH_TREE_QUADTREE || c->nsys)
        return fail(c, BH_ERR_STATE, "the learned correction needs a one-system QUADTREE context");
    if (enable && !c->m2) {
        bh_status st = dalloc(c, &c->m2, c->cap_nodes);
        if (st == BH_OK) st = dalloc(c, &c->lfeat, (size_t)c->cap_nodes * kMlpIn);
        if (st == BH_OK) st = dalloc(c, &c->lcoef, (size_t)c->cap_nodes * (kCoefStride / 4));
        if (st == BH_OK) st = dalloc(c, &c->lbig, cell_features_lists_words(c->cfg.n_max));
        if (st != BH_OK) return st;
    }
    c->learned = enable != 0;
    c->lfeat_ok = false;
    c->built = false;  // second moments appear with the next build
    return BH_OK;
}

bh_status bh_learned_collect(bh_ctx* c, bh_mlp* m, double theta, double band,
                             const uint32_t* sinks, uint32_t ns, uint32_t max_pairs_per_sink,
                             uint32_t* added) {
    if (!c || !m || (!sinks && ns) || !(theta > 0.0) || band < 0.0 || max_pairs_per_sink == 0)
        return BH_ERR_ARG;
    if (!c->learned || !c->built) return fail(c, BH_ERR_STATE, "enable the learned mode and build first");
    if (ensure_features(c) != BH_OK) return BH_ERR_HIP;
    if (ns == 0) return BH_OK;
    for (uint32_t k = 0; k < ns; ++k)
        if (sinks[k] >= c->n) return fail(c, BH_ERR_ARG, "sink index %u out of range", sinks[k]);
    hipStream_t s = c->stream;
    uint32_t *d_sc = nullptr, *d_ss = nullptr, *pairs = nullptr, *counts = nullptr, *offs = nullptr,
             *spart = nullptr;
    const size_t np = (size_t)ns * max_pairs_per_sink;
    bh_status st = hip_check(c, hipMalloc((void**)&d_sc, (size_t)ns * 4), "hipMalloc");
    if (st == BH_OK) st = hip_check(c, hipMalloc((void**)&d_ss, (size_t)ns * 4), "hipMalloc");
    if (st == BH_OK) st = hip_check(c, hipMalloc((void**)&pairs, np * 4), "hipMalloc");
    if (st == BH_OK) st = hip_check(c, hipMalloc((void**)&counts, (size_t)ns * 4), "hipMalloc");
    if (st == BH_OK) st = hip_check(c, hipMalloc((void**)&offs, (size_t)ns * 4), "hipMalloc");
    const size_t nsp = scan_partials_size(ns);
    if (st == BH_OK) st = hip_check(c, hipMalloc((void**)&spart, (nsp + 1) * 4), "hipMalloc");
    if (st == BH_OK) st = hip_check(c, hipMemcpyAsync(d_sc, sinks, (size_t)ns * 4, hipMemcpyHostToDevice, s), "sinks");
    uint32_t last[1] = {0};
    if (st == BH_OK) {
        launch_sinks_sorted(s, c->id[c->cur], c->n, d_sc, ns, c->where, d_ss);
        const float inv_t2 = (float)(1.0 / (theta * theta));
        const float inv_b2 = band > 0.0 ? (float)(1.0 / (band * band)) : 0.f;
        launch_collect_pairs(s, c->pos4, c->nodes, c->prm, d_ss, ns, inv_t2, inv_b2,
                             max_pairs_per_sink, pairs, counts);
        launch_exclusive_scan(s, counts, offs, ns, spart, spart + nsp);
        st = launch_status(c, "collect");
    }
    if (st == BH_OK) st = hip_check(c, hipMemcpyAsync(&last[0], spart + nsp, 4, hipMemcpyDeviceToHost, s), "total");
    if (st == BH_OK) st = hip_check(c, hipStreamSynchronize(s), "collect");
    const uint64_t total = last[0];
    if (st == BH_OK && m->n + total > m->cap)
        st = fail(c, BH_ERR_CAPACITY, "data set capacity %u exceeded (%u + %llu)", m->cap, m->n,
                  (unsigned long long)total);
    if (st == BH_OK) {
        const double eps = c->cfg.softening;
        launch_label_pairs(s, c->pos4, c->nodes, c->lfeat, c->prm, d_ss, pairs, counts, offs, ns,
                           max_pairs_per_sink, eps * eps, m->X + (size_t)m->n * kLearnSample,
                           m->T + (size_t)m->n * 2, (uint32_t)total);
        st = launch_status(c, "label");
    }
    if (st == BH_OK) st = hip_check(c, hipStreamSynchronize(s), "label");
    (void)hipStreamSynchronize(s);
    for (uint32_t* p : {d_sc, d_ss, pairs, counts, offs, spart})
        if (p) (void)hipFree(p);
    if (st == BH_OK) {
        m->n += (uint32_t)total;
        if (added) *added = (uint32_t)total;
    }
    return st;
}

bh_status bh_forces_learned(bh_ctx* c, const bh_mlp* m, double theta, double band, float* ax,
                            float* ay, double* total) {
    if (!c || !m || !(theta > 0.0) || band < 0.0) return BH_ERR_ARG;
    if (!c->learned || !c->built) return fail(c, BH_ERR_STATE, "enable the learned mode and build first");
    if (ensure_features(c) != BH_OK) return BH_ERR_HIP;
    hipStream_t s = c->stream;
    const float inv_t2 = (float)(1.0 / (theta * theta));
    const float inv_b2 = band > 0.0 ? (float)(1.0 / (band * band)) : 0.f;
    const float eps = c->cfg.softening;
    phase_begin(c, 5);  // the per-cell inference is part of the learned force evaluation
    launch_cell_coefs(s, c->lfeat, c->nodes, c->prm, c->cap_nodes, m->wbf, c->lcoef);
    c->trav_blocks = launch_traverse_learned(s, c->variant, c->pos4, c->nodes, c->lcoef, c->prm,
                                             c->n, inv_t2, inv_b2, eps * eps, c->cfg.G, m->terms,
                                             c->ax, c->ay, total ? c->tot : nullptr, c->counters);
    phase_end(c, 5);
    if (total) launch_reduce_f64(s, c->tot, nullptr, c->n, c->red_part, c->dsum);
    c->have_acc = true;
    bh_status st = launch_status(c, "learned traverse");
    if (st != BH_OK) return st;
    const uint32_t n = c->n;
    if (ax || ay) {
        launch_unsort2(s, c->id[c->cur], n, c->ax, c->ay, c->out_a, c->out_b);
        if (ax) st = hip_check(c, hipMemcpyAsync(ax, c->out_a, (size_t)n * 4, hipMemcpyDeviceToHost, s), "ax");
        if (st == BH_OK && ay) st = hip_check(c, hipMemcpyAsync(ay, c->out_b, (size_t)n * 4, hipMemcpyDeviceToHost, s), "ay");
    }
    if (st == BH_OK && total) st = hip_check(c, hipMemcpyAsync(c->h_dsum, c->dsum, 8, hipMemcpyDeviceToHost, s), "total");
    if (st == BH_OK) st = hip_check(c, hipStreamSynchronize(s), "learned forces");
    if (st == BH_OK && total) *total = *c->h_dsum;
    return st;
}

bh_status bh_learned_export(bh_ctx* c, const bh_mlp* m, float* features, float* coef,
                            uint32_t n_nodes) {
    if (!c || (!features && !coef) || (coef && !m)) return BH_ERR_ARG;
    if (!c->learned || !c->built) return fail(c, BH_ERR_STATE, "enable the learned mode and build first");
    hipStream_t s = c->stream;
    bh_status st = coef ? ensure_coefs(c, m) : ensure_features(c);
    uint32_t nn = 0;
    if (st == BH_OK) st = hip_check(c, hipMemcpyAsync(&c->h_prm->n_nodes, &c->prm->n_nodes, 4, hipMemcpyDeviceToHost, s), "nodes");
    if (st == BH_OK) st = hip_check(c, hipStreamSynchronize(s), "nodes");
    if (st == BH_OK) nn = c->h_prm->n_nodes;
    if (st == BH_OK && n_nodes != nn)
        return fail(c, BH_ERR_ARG, "the tree has %u nodes, %u requested", nn, n_nodes);
    if (st == BH_OK && features)
        st = hip_check(c, hipMemcpyAsync(features, c->lfeat, (size_t)nn * kMlpIn * 4, hipMemcpyDeviceToHost, s), "features");
    if (st == BH_OK && coef)
        st = hip_check(c, hipMemcpyAsync(coef, c->lcoef, (size_t)nn * kCoefStride * 4, hipMemcpyDeviceToHost, s), "coefficients");
    if (st == BH_OK) st = hip_check(c, hipStreamSynchronize(s), "export");
    return st;
}

}  // extern "C"
