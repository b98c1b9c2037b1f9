/*
 * CPU oracle for the Barnes-Hut force step — TEST INFRASTRUCTURE ONLY (see bh_oracle.h).
 * Plain C restatement of the reference algorithm; FP64 like the reference. Not the product.
 */
#define _GNU_SOURCE
#include "bh_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------------------------ */
/* Morton codes                                                                                */
/* ------------------------------------------------------------------------------------------ */

/* bit spread of a 16-bit value into the even bits (nBodySolverFMM_SYM_MORTON.h:554-561) */
static uint32_t spread16(uint32_t x) {
    x &= 0xFFFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

static uint64_t spread32(uint64_t x) {
    x &= 0xFFFFFFFFull;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

/* glm::min / glm::max scalar forms used by the reference: max(x,y) = x<y ? y : x */
static double gmax(double x, double y) { return (x < y) ? y : x; }
static double gmin(double x, double y) { return (y < x) ? y : x; }

/* createMortonCode (nBodySolverFMM_SYM_MORTON.h:538-552): t = (p + (-min)) / L, scaled by
 * 65536, clamped to [0, 65535], truncated; code = spread(y) << 1 | spread(x). */
void orc_morton32(const double* x, const double* y, int64_t n, double minx, double miny, double L,
                  uint32_t* code) {
    const double nmx = -minx, nmy = -miny;
    for (int64_t i = 0; i < n; ++i) {
        double px = (x[i] + nmx) / L;
        double py = (y[i] + nmy) / L;
        px = gmin(gmax(px * 65536.0, 0.0), 65535.0);
        py = gmin(gmax(py * 65536.0, 0.0), 65535.0);
        code[i] = (spread16((uint32_t)py) << 1) | spread16((uint32_t)px);
    }
}

/* createRoot (nBodySolverGpuBH.cu:78-94): centre = FP32 midpoint of the bounding box, side =
 * the larger FP32 extent. createChildren (:96-200): a cell's points are split by x against the
 * centre (<= goes low), each half by y; the four children, in the order (x, y) = (low, low),
 * (low, high), (high, low), (high, high), get centres c + (+-0.25 l, +-0.25 l) and side 0.5 l in
 * FP32. Restated as the reference does it, cell by cell over index lists, level by level. */
void orc_gpuref_codes32(const float* x, const float* y, int64_t n, int depth, uint32_t* code) {
    if (n <= 0) return;
    float mnx = x[0], mny = y[0], mxx = x[0], mxy = y[0];
    for (int64_t i = 1; i < n; ++i) {
        if (x[i] < mnx) mnx = x[i];
        if (y[i] < mny) mny = y[i];
        if (x[i] > mxx) mxx = x[i];
        if (y[i] > mxy) mxy = y[i];
    }
    const float lx = mxx - mnx, ly = mxy - mny;
    /* cells of the current level: point lists (as index ranges of `idx`), centres, sides */
    int64_t ncell = 1;
    int64_t* idx = (int64_t*)malloc((size_t)n * sizeof(int64_t));
    int64_t* tmp = (int64_t*)malloc((size_t)n * sizeof(int64_t));
    int64_t* beg = (int64_t*)malloc(sizeof(int64_t) * 2);
    float* cx = (float*)malloc(sizeof(float));
    float* cy = (float*)malloc(sizeof(float));
    float* len = (float*)malloc(sizeof(float));
    for (int64_t i = 0; i < n; ++i) {
        idx[i] = i;
        code[i] = 0;
    }
    beg[0] = 0;
    beg[1] = n;
    cx[0] = (mnx + mxx) * 0.5f;
    cy[0] = (mny + mxy) * 0.5f;
    len[0] = lx > ly ? lx : ly;
    for (int l = 0; l < depth; ++l) {
        const int64_t nc = ncell * 4;
        int64_t* nbeg = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nc + 1));
        float* ncx = (float*)malloc(sizeof(float) * (size_t)nc);
        float* ncy = (float*)malloc(sizeof(float) * (size_t)nc);
        float* nlen = (float*)malloc(sizeof(float) * (size_t)nc);
        int64_t w = 0;
        for (int64_t c = 0; c < ncell; ++c) {
            for (int q = 0; q < 4; ++q) { /* A, B, C, D */
                const int xh = q >> 1, yh = q & 1;
                nbeg[4 * c + q] = w;
                for (int64_t k = beg[c]; k < beg[c + 1]; ++k) {
                    const int64_t i = idx[k];
                    const int px = !(x[i] <= cx[c]), py = !(y[i] <= cy[c]);
                    if (px == xh && py == yh) {
                        tmp[w++] = i;
                        code[i] |= (uint32_t)((yh << 1) | xh) << (30 - 2 * l);
                    }
                }
                const float o = 0.25f * len[c];
                ncx[4 * c + q] = cx[c] + (xh ? o : -o);
                ncy[4 * c + q] = cy[c] + (yh ? o : -o);
                nlen[4 * c + q] = 0.5f * len[c];
            }
        }
        nbeg[nc] = w;
        int64_t* s = idx;
        idx = tmp;
        tmp = s;
        free(beg);
        free(cx);
        free(cy);
        free(len);
        beg = nbeg;
        cx = ncx;
        cy = ncy;
        len = nlen;
        ncell = nc;
    }
    free(idx);
    free(tmp);
    free(beg);
    free(cx);
    free(cy);
    free(len);
}

void orc_quantize_axis(const double* v, int64_t n, double vmin, double L, int bits, uint64_t* q) {
    const double nm = -vmin;
    const double scale = ldexp(1.0, bits);
    for (int64_t i = 0; i < n; ++i) {
        double t = (v[i] + nm) / L;
        t = gmin(gmax(t * scale, 0.0), scale - 1.0);
        q[i] = (uint64_t)t;
    }
}

/* Same rule as createMortonCode with `bits` bits per axis (bits <= 32). */
void orc_morton64(const double* x, const double* y, int64_t n, double minx, double miny, double L,
                  int bits, uint64_t* code) {
    const double nmx = -minx, nmy = -miny;
    const double scale = ldexp(1.0, bits);
    for (int64_t i = 0; i < n; ++i) {
        double px = (x[i] + nmx) / L;
        double py = (y[i] + nmy) / L;
        px = gmin(gmax(px * scale, 0.0), scale - 1.0);
        py = gmin(gmax(py * scale, 0.0), scale - 1.0);
        code[i] = (spread32((uint64_t)py) << 1) | spread32((uint64_t)px);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* stable sort: LSD radix on 16-bit digits, ties keep the original index order.                */
/* The reference's boost integer_sort (nBodySolverFMM_SYM_MORTON.h:89) leaves tie order        */
/* unspecified; the tie rule here (original index) is this project's definition.              */
/* ------------------------------------------------------------------------------------------ */
void orc_stable_sort_u64(const uint64_t* key, int64_t n, uint32_t* perm) {
    uint32_t* tmp = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    uint64_t* cnt = (uint64_t*)malloc(sizeof(uint64_t) * 65536);
    for (int64_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
    uint32_t* src = perm;
    uint32_t* dst = tmp;
    for (int pass = 0; pass < 4; ++pass) {
        const int sh = 16 * pass;
        memset(cnt, 0, sizeof(uint64_t) * 65536);
        for (int64_t i = 0; i < n; ++i) cnt[(key[src[i]] >> sh) & 0xFFFF]++;
        int trivial = 0;
        for (int d = 0; d < 65536; ++d)
            if (cnt[d] == (uint64_t)n) trivial = 1;
        if (trivial) continue;
        uint64_t run = 0;
        for (int d = 0; d < 65536; ++d) {
            uint64_t c = cnt[d];
            cnt[d] = run;
            run += c;
        }
        for (int64_t i = 0; i < n; ++i) dst[cnt[(key[src[i]] >> sh) & 0xFFFF]++] = src[i];
        uint32_t* t = src;
        src = dst;
        dst = t;
    }
    if (src != perm) memcpy(perm, src, sizeof(uint32_t) * (size_t)n);
    free(tmp);
    free(cnt);
}

/* ------------------------------------------------------------------------------------------ */
/* fixed-depth Morton tree                                                                      */
/* ------------------------------------------------------------------------------------------ */

/* getDepth (nBodySolverFMM_SYM_MORTON.h:106-115). N/max_children is an int division. For
 * N < max_children the reference takes log(0) and casts -inf to unsigned (undefined); this
 * restatement returns 0 there (root only). */
int orc_get_depth(int max_children, int64_t n) {
    int q = (int)n / max_children;
    if (q <= 0) return 0;
    double circle_area = M_PI / 4.0;
    double depth_real = log((1.0 / circle_area) * (double)q) / log(4.0);
    double c = ceil(depth_real);
    return c < 0 ? 0 : (int)c;
}

/* initNodesSize (nBodySolverFMM_SYM_MORTON.h:117-151): levelIndex[l] = sum_{j<l} 4^j */
static int64_t level_index(int l) { return (((int64_t)1 << (2 * l)) - 1) / 3; }
int64_t orc_fixed_tree_slots(int depth) { return level_index(depth + 1); }

/* createLeafNodes + bottomUpNodeConstruction (nBodySolverFMM_SYM_MORTON.h:328-470), over points
 * already sorted by Morton code. m2 holds 3 doubles per slot (xx, xy, yy). */
void orc_fixed_tree(const double* xs, const double* ys, int64_t n, int depth, double minx,
                    double miny, double maxx, double maxy, uint32_t* first_child,
                    uint32_t* first_particle, uint32_t* count, double* bb_cx, double* bb_cy,
                    double* bb_len, double* com_x, double* com_y, double* m0, double* m2) {
    const int64_t slots = orc_fixed_tree_slots(depth);
    memset(first_child, 0, sizeof(uint32_t) * slots);
    memset(first_particle, 0, sizeof(uint32_t) * slots);
    memset(count, 0, sizeof(uint32_t) * slots);
    memset(bb_cx, 0, sizeof(double) * slots);
    memset(bb_cy, 0, sizeof(double) * slots);
    memset(bb_len, 0, sizeof(double) * slots);
    memset(com_x, 0, sizeof(double) * slots);
    memset(com_y, 0, sizeof(double) * slots);
    memset(m0, 0, sizeof(double) * slots);
    memset(m2, 0, sizeof(double) * 3 * slots);

    const double nmx = -minx, nmy = -miny;
    const double largest = gmax(maxx + nmx, maxy + nmy); /* glm::compMax(maxPos + negMinPos) */
    const uint32_t W = 1u << depth;
    const double leaf_size = largest / (double)W;
    const int64_t li = level_index(depth);

    for (int64_t i = 0; i < n; ++i) {
        double gx = ((double)W * (xs[i] + nmx)) / largest;
        double gy = ((double)W * (ys[i] + nmy)) / largest;
        uint32_t cx = (uint32_t)gx, cy = (uint32_t)gy;
        if (cx > W - 1) cx = W - 1;
        if (cy > W - 1) cy = W - 1;
        int64_t leaf = li + ((spread16(cy) << 1) | spread16(cx));
        if (count[leaf] == 0u) {
            bb_cx[leaf] = (minx + leaf_size * (double)cx) + 0.5 * leaf_size;
            bb_cy[leaf] = (miny + leaf_size * (double)cy) + 0.5 * leaf_size;
            bb_len[leaf] = leaf_size;
            first_particle[leaf] = (uint32_t)i;
        }
        count[leaf] += 1u;
        com_x[leaf] += xs[i];
        com_y[leaf] += ys[i];
    }
    for (int64_t s = li; s < li + ((int64_t)1 << (2 * depth)); ++s) {
        if (count[s] == 0u) continue;
        m0[s] = (double)count[s];
        com_x[s] /= m0[s];
        com_y[s] /= m0[s];
        for (int64_t p = first_particle[s]; p < (int64_t)first_particle[s] + count[s]; ++p) {
            double rx = xs[p] - com_x[s], ry = ys[p] - com_y[s];
            m2[3 * s + 0] += rx * rx;
            m2[3 * s + 1] += rx * ry;
            m2[3 * s + 2] += ry * ry;
        }
    }
    static const double off[4][2] = {{0.5, 0.5}, {-0.5, 0.5}, {0.5, -0.5}, {-0.5, -0.5}};
    for (int l = depth - 1; l >= 0; --l) {
        const int64_t size = (int64_t)1 << (2 * l);
        for (int64_t i = 0; i < size; ++i) {
            const int64_t node = level_index(l) + i;
            const int64_t pfc = level_index(l + 1) + 4 * i;
            for (int j = 3; j >= 0; --j) {
                const int64_t c = pfc + j;
                if (count[c] == 0u) continue;
                first_child[node] = (uint32_t)pfc;
                bb_cx[node] = off[j][0] * bb_len[c] + bb_cx[c];
                bb_cy[node] = off[j][1] * bb_len[c] + bb_cy[c];
                bb_len[node] = 2.0 * bb_len[c];
                first_particle[node] = first_particle[c];
                count[node] += count[c];
                com_x[node] += m0[c] * com_x[c];
                com_y[node] += m0[c] * com_y[c];
                m0[node] += m0[c];
            }
            if (count[node] == 0u) continue;
            com_x[node] /= m0[node];
            com_y[node] /= m0[node];
            for (int64_t c = first_child[node]; c < (int64_t)first_child[node] + 4; ++c) {
                if (m0[c] == 0.0) continue;
                double rx = com_x[c] - com_x[node], ry = com_y[c] - com_y[node];
                /* the reference adds the outer product first, then the child's M2, and does
                 * not weight the outer product by the child's M0 (:451-466) */
                m2[3 * node + 0] += rx * rx;
                m2[3 * node + 1] += rx * ry;
                m2[3 * node + 2] += ry * ry;
                m2[3 * node + 0] += m2[3 * c + 0];
                m2[3 * node + 1] += m2[3 * c + 1];
                m2[3 * node + 2] += m2[3 * c + 2];
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* pre-order trees over sorted codes                                                            */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    const uint64_t* codes64;
    const uint32_t* codes32;
    int bits; /* bits per axis of the code */
    int leaf_cap;
    int fixed_depth; /* >= 0: fixed-depth mode */
    int64_t cap_nodes, nn;
    int32_t* level;
    uint32_t *first, *count, *skip, *nchild;
    int overflow;
} pre_ctx;

static uint64_t code_at(const pre_ctx* c, int64_t i) {
    return c->codes64 ? c->codes64[i] : (uint64_t)c->codes32[i];
}
/* digit of level k (k >= 1) */
static int digit_at(const pre_ctx* c, int64_t i, int k) {
    return (int)((code_at(c, i) >> (2 * (c->bits - k))) & 3u);
}

static void pre_build(pre_ctx* c, int64_t first, int64_t count, int lvl) {
    if (c->nn >= c->cap_nodes) {
        c->overflow = 1;
        return;
    }
    const int64_t me = c->nn++;
    c->level[me] = lvl;
    c->first[me] = (uint32_t)first;
    c->count[me] = (uint32_t)count;
    c->nchild[me] = 0;
    int split;
    if (c->fixed_depth >= 0)
        split = lvl < c->fixed_depth;
    else
        split = count > c->leaf_cap && lvl < c->bits;
    if (split) {
        int64_t s = first;
        for (int d = 0; d < 4; ++d) {
            int64_t e = s;
            while (e < first + count && digit_at(c, e, lvl + 1) == d) ++e;
            if (e > s) {
                c->nchild[me]++;
                pre_build(c, s, e - s, lvl + 1);
            }
            s = e;
        }
    }
    c->skip[me] = (uint32_t)c->nn;
}

int64_t orc_adaptive_tree(const uint64_t* codes_sorted, int64_t n, int bits, int leaf_cap,
                          int64_t cap_nodes, int32_t* level, uint32_t* first, uint32_t* count,
                          uint32_t* skip, uint32_t* nchild) {
    pre_ctx c = {codes_sorted, NULL, bits, leaf_cap, -1, cap_nodes, 0, level, first, count,
                 skip, nchild, 0};
    if (n <= 0) return 0;
    pre_build(&c, 0, n, 0);
    return c.overflow ? -1 : c.nn;
}

int64_t orc_fixed_preorder(const uint32_t* codes_sorted, int64_t n, int depth, int64_t cap_nodes,
                           int32_t* level, uint32_t* first, uint32_t* count, uint32_t* skip,
                           uint32_t* nchild) {
    pre_ctx c = {NULL, codes_sorted, 16, 0, depth, cap_nodes, 0, level, first, count,
                 skip, nchild, 0};
    if (n <= 0) return 0;
    pre_build(&c, 0, n, 0);
    return c.overflow ? -1 : c.nn;
}

void orc_node_moments(const double* xs, const double* ys, const double* ms, int64_t nnodes,
                      const uint32_t* first, const uint32_t* count, double* cx, double* cy,
                      double* mass, double* m2) {
    for (int64_t k = 0; k < nnodes; ++k) {
        double sm = 0, sx = 0, sy = 0;
        for (int64_t p = first[k]; p < (int64_t)first[k] + count[k]; ++p) {
            double mp = ms ? ms[p] : 1.0;
            sm += mp;
            sx += mp * xs[p];
            sy += mp * ys[p];
        }
        mass[k] = sm;
        cx[k] = sx / sm;
        cy[k] = sy / sm;
        if (m2) {
            double a = 0, b = 0, d = 0;
            for (int64_t p = first[k]; p < (int64_t)first[k] + count[k]; ++p) {
                double mp = ms ? ms[p] : 1.0;
                double rx = xs[p] - cx[k], ry = ys[p] - cy[k];
                a += mp * rx * rx;
                b += mp * rx * ry;
                d += mp * ry * ry;
            }
            m2[3 * k + 0] = a;
            m2[3 * k + 1] = b;
            m2[3 * k + 2] = d;
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* interaction kernels                                                                          */
/* ------------------------------------------------------------------------------------------ */

/* t-SNE PN/PP kernels (nBodySolverBH.h:94-116): w = 1/(1+|p-c|^2); total += M w;
 * derivative += M w^2 (p - c). Gravity: a += G M (c - p) / (|c-p|^2 + eps^2)^(3/2);
 * total += -G M / sqrt(|c-p|^2 + eps^2) (the sink's potential). */
static inline void interact(int kernel, double G, double eps2, double px, double py, double cx,
                            double cy, double M, double* ax, double* ay, double* tot) {
    if (kernel == ORC_KERNEL_TSNE) {
        double dx = px - cx, dy = py - cy;
        double sq = dx * dx + dy * dy;
        double fd = 1.0 / (1.0 + sq);
        *tot += M * fd;
        double s = M * fd * fd;
        *ax += s * dx;
        *ay += s * dy;
    } else {
        double dx = cx - px, dy = cy - py;
        double r2 = dx * dx + dy * dy + eps2;
        double inv = 1.0 / sqrt(r2);
        double s = G * M * inv * inv * inv;
        *ax += s * dx;
        *ay += s * dy;
        *tot += -G * M * inv;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* BH traversal over a pre-order tree                                                           */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    const double *xs, *ys, *ms;
    int64_t n;
    const int32_t* level;
    const uint32_t *first, *count, *skip;
    const uint8_t* pp;
    const double *cx, *cy, *mass;
    int64_t nnodes;
    double root_side;
    int semantics, kernel;
    double G, eps2, theta;
    const int64_t* sinks;
    int64_t nsinks;
    double *ax, *ay, *tot;
    int64_t lo, hi;
    uint64_t pn, ppc;
} trav_job;

static void* trav_worker(void* arg) {
    trav_job* j = (trav_job*)arg;
    uint64_t pn = 0, ppc = 0;
    for (int64_t s = j->lo; s < j->hi; ++s) {
        const int64_t i = j->sinks ? j->sinks[s] : s;
        const double px = j->xs[i], py = j->ys[i];
        double ax = 0, ay = 0, tot = 0;
        int64_t k = 0;
        while (k < j->nnodes) {
            const double l = ldexp(j->root_side, -j->level[k]);
            const double dx = px - j->cx[k], dy = py - j->cy[k];
            const double d2 = dx * dx + dy * dy;
            int accept;
            if (j->semantics == ORC_SEM_QUADTREE) {
                accept = (l / sqrt(d2)) < j->theta;
            } else {
                const double thr = l / j->theta;
                accept = d2 > thr * thr;
            }
            if (accept) {
                interact(j->kernel, j->G, j->eps2, px, py, j->cx[k], j->cy[k], j->mass[k], &ax,
                         &ay, &tot);
                ++pn;
                k = j->skip[k];
            } else if (j->pp[k]) {
                const int64_t f = j->first[k], e = f + j->count[k];
                for (int64_t q = f; q < e; ++q) {
                    if (q == i) continue;
                    interact(j->kernel, j->G, j->eps2, px, py, j->xs[q], j->ys[q],
                             j->ms ? j->ms[q] : 1.0, &ax, &ay, &tot);
                    ++ppc;
                }
                k = j->skip[k];
            } else {
                ++k;
            }
        }
        j->ax[s] = ax;
        j->ay[s] = ay;
        if (j->tot) j->tot[s] = tot;
    }
    j->pn = pn;
    j->ppc = ppc;
    return NULL;
}

void orc_bh_traverse(const double* xs, const double* ys, const double* ms, int64_t n,
                     const int32_t* level, const uint32_t* first, const uint32_t* count,
                     const uint32_t* skip, const uint8_t* pp, const double* cx, const double* cy,
                     const double* mass, int64_t nnodes, double root_side, int semantics,
                     int kernel, double G, double eps2, double theta, const int64_t* sinks,
                     int64_t nsinks, double* out_ax, double* out_ay, double* out_tot,
                     uint64_t* counters, int nthreads) {
    if (!sinks) nsinks = n;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    trav_job jobs[256];
    pthread_t th[256];
    const int64_t chunk = (nsinks + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        trav_job* j = &jobs[t];
        j->xs = xs; j->ys = ys; j->ms = ms; j->n = n;
        j->level = level; j->first = first; j->count = count; j->skip = skip; j->pp = pp;
        j->cx = cx; j->cy = cy; j->mass = mass; j->nnodes = nnodes; j->root_side = root_side;
        j->semantics = semantics; j->kernel = kernel; j->G = G; j->eps2 = eps2; j->theta = theta;
        j->sinks = sinks; j->nsinks = nsinks; j->ax = out_ax; j->ay = out_ay; j->tot = out_tot;
        j->lo = t * chunk;
        j->hi = (t + 1) * chunk < nsinks ? (t + 1) * chunk : nsinks;
        if (j->lo > j->hi) j->lo = j->hi;
        j->pn = j->ppc = 0;
    }
    if (nthreads == 1) {
        trav_worker(&jobs[0]);
    } else {
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, trav_worker, &jobs[t]);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    }
    if (counters) {
        counters[0] = counters[1] = 0;
        for (int t = 0; t < nthreads; ++t) {
            counters[0] += jobs[t].pn;
            counters[1] += jobs[t].ppc;
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Barnes-Hut walk with single-precision acceptance decisions on given node records            */
/* ------------------------------------------------------------------------------------------ */
/* The walk of orc_bh_traverse (nBodySolverBH.h:56-89 control flow: accept -> monopole and skip
 * the subtree, else a near-field cell -> particle-particle over its range, else descend), with
 * the acceptance taken as the single-precision test d2 > l2 * (1/theta^2):
 *   l = root_side * 2^-level (exact), l2 = RN32(l * l), lim = RN32(l2 * inv_theta2),
 *   dx = RN32(cx - px), dy = RN32(cy - py), d2 = fma32(dx, dx, RN32(dy * dy)).
 * Every interaction itself is evaluated in FP64 from the same single-precision inputs, so the
 * result is the exact sum of the interactions the single-precision decisions select; abs_x/abs_y
 * receive the per-sink sums of |contribution| that bound the rounding of any evaluation order.
 * With FP64 node moments given (cx64 != NULL), each decision on the walk is also taken by the
 * reference's FP64 rule of `semantics` (QUADTREE: l/|d| < theta, nBodySolverBH.h:58-62; MORTON:
 * d^2 > (l/theta)^2, nBodySolverFMM_SYM_MORTON.h:276-280) and disagreements are counted. */
typedef struct {
    const float *xs, *ys, *ms;
    const int32_t* level;
    const uint32_t *first, *count, *skip;
    const uint8_t* pp;
    const float *cx, *cy, *mass;
    int64_t nnodes;
    float root_side, inv_theta2;
    const double *cx64, *cy64;
    double side64, theta;
    int semantics, kernel;
    double G, eps2;
    const int64_t* sinks;
    double *ax, *ay, *tot, *abs_x, *abs_y;
    int64_t lo, hi;
    uint64_t pn, ppc, mism;
} trav32_job;

static void* trav32_worker(void* arg) {
    trav32_job* j = (trav32_job*)arg;
    uint64_t pn = 0, ppc = 0, mism = 0;
    for (int64_t s = j->lo; s < j->hi; ++s) {
        const int64_t i = j->sinks ? j->sinks[s] : s;
        const float px = j->xs[i], py = j->ys[i];
        double ax = 0, ay = 0, tot = 0, bx = 0, by = 0;
        int64_t k = 0;
        while (k < j->nnodes) {
            const float sl = ldexpf(j->root_side, -j->level[k]);
            const float l2 = sl * sl;
            const float lim = l2 * j->inv_theta2;
            const float dx = j->cx[k] - px, dy = j->cy[k] - py;
            const float d2 = fmaf(dx, dx, dy * dy);
            const int accept = d2 > lim;
            if (j->cx64) {
                const double l = ldexp(j->side64, -j->level[k]);
                const double ex = (double)px - j->cx64[k], ey = (double)py - j->cy64[k];
                const double e2 = ex * ex + ey * ey;
                int acc64;
                if (j->semantics == ORC_SEM_QUADTREE) {
                    acc64 = (l / sqrt(e2)) < j->theta;
                } else {
                    const double thr = l / j->theta;
                    acc64 = e2 > thr * thr;
                }
                mism += (uint64_t)(acc64 != accept);
            }
            if (accept) {
                double cxv = 0, cyv = 0, ctv = 0;
                interact(j->kernel, j->G, j->eps2, px, py, j->cx[k], j->cy[k], j->mass[k], &cxv,
                         &cyv, &ctv);
                ax += cxv; ay += cyv; tot += ctv;
                bx += fabs(cxv); by += fabs(cyv);
                ++pn;
                k = j->skip[k];
            } else if (j->pp[k]) {
                const int64_t f = j->first[k], e = f + j->count[k];
                for (int64_t q = f; q < e; ++q) {
                    if (q == i) continue;
                    double cxv = 0, cyv = 0, ctv = 0;
                    interact(j->kernel, j->G, j->eps2, px, py, j->xs[q], j->ys[q],
                             j->ms ? j->ms[q] : 1.0, &cxv, &cyv, &ctv);
                    ax += cxv; ay += cyv; tot += ctv;
                    bx += fabs(cxv); by += fabs(cyv);
                    ++ppc;
                }
                k = j->skip[k];
            } else {
                ++k;
            }
        }
        j->ax[s] = ax;
        j->ay[s] = ay;
        if (j->tot) j->tot[s] = tot;
        if (j->abs_x) j->abs_x[s] = bx;
        if (j->abs_y) j->abs_y[s] = by;
    }
    j->pn = pn;
    j->ppc = ppc;
    j->mism = mism;
    return NULL;
}

void orc_bh_traverse_f32(const float* xs, const float* ys, const float* ms, int64_t n,
                         const int32_t* level, const uint32_t* first, const uint32_t* count,
                         const uint32_t* skip, const uint8_t* pp, const float* cx, const float* cy,
                         const float* mass, int64_t nnodes, float root_side, float inv_theta2,
                         int kernel, double G, double eps2, const double* cx64, const double* cy64,
                         double side64, double theta, int semantics, const int64_t* sinks,
                         int64_t nsinks, double* out_ax, double* out_ay, double* out_tot,
                         double* abs_x, double* abs_y, uint64_t* counters, int nthreads) {
    if (!sinks) nsinks = n;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    trav32_job jobs[256];
    pthread_t th[256];
    const int64_t chunk = (nsinks + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        trav32_job* j = &jobs[t];
        memset(j, 0, sizeof(*j));
        j->xs = xs; j->ys = ys; j->ms = ms;
        j->level = level; j->first = first; j->count = count; j->skip = skip; j->pp = pp;
        j->cx = cx; j->cy = cy; j->mass = mass; j->nnodes = nnodes;
        j->root_side = root_side; j->inv_theta2 = inv_theta2;
        j->cx64 = cx64; j->cy64 = cy64; j->side64 = side64; j->theta = theta;
        j->semantics = semantics; j->kernel = kernel; j->G = G; j->eps2 = eps2;
        j->sinks = sinks; j->ax = out_ax; j->ay = out_ay; j->tot = out_tot;
        j->abs_x = abs_x; j->abs_y = abs_y;
        j->lo = t * chunk;
        j->hi = (t + 1) * chunk < nsinks ? (t + 1) * chunk : nsinks;
        if (j->lo > j->hi) j->lo = j->hi;
    }
    (void)n;
    if (nthreads == 1) {
        trav32_worker(&jobs[0]);
    } else {
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, trav32_worker, &jobs[t]);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    }
    if (counters) {
        counters[0] = counters[1] = counters[2] = 0;
        for (int t = 0; t < nthreads; ++t) {
            counters[0] += jobs[t].pn;
            counters[1] += jobs[t].ppc;
            counters[2] += jobs[t].mism;
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* FP64 compensated direct sum                                                                  */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    const double *x, *y, *m;
    int64_t n;
    int kernel;
    double G, eps2;
    const int64_t* sinks;
    double *ax, *ay, *tot;
    int64_t lo, hi;
} direct_job;

static inline void kahan_add(double* s, double* c, double v) {
    double yv = v - *c;
    double t = *s + yv;
    *c = (t - *s) - yv;
    *s = t;
}

static void* direct_worker(void* arg) {
    direct_job* j = (direct_job*)arg;
    for (int64_t s = j->lo; s < j->hi; ++s) {
        const int64_t i = j->sinks ? j->sinks[s] : s;
        const double px = j->x[i], py = j->y[i];
        double sx = 0, cx = 0, sy = 0, cy = 0, st = 0, ct = 0;
        for (int64_t q = 0; q < j->n; ++q) {
            if (q == i) continue;
            double ax = 0, ay = 0, tot = 0;
            interact(j->kernel, j->G, j->eps2, px, py, j->x[q], j->y[q], j->m ? j->m[q] : 1.0,
                     &ax, &ay, &tot);
            kahan_add(&sx, &cx, ax);
            kahan_add(&sy, &cy, ay);
            kahan_add(&st, &ct, tot);
        }
        j->ax[s] = sx;
        j->ay[s] = sy;
        if (j->tot) j->tot[s] = st;
    }
    return NULL;
}

void orc_direct(const double* x, const double* y, const double* m, int64_t n, int kernel, double G,
                double eps2, const int64_t* sinks, int64_t nsinks, double* out_ax, double* out_ay,
                double* out_tot, int nthreads) {
    if (!sinks) nsinks = n;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    direct_job jobs[256];
    pthread_t th[256];
    const int64_t chunk = (nsinks + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        direct_job* j = &jobs[t];
        j->x = x; j->y = y; j->m = m; j->n = n; j->kernel = kernel; j->G = G; j->eps2 = eps2;
        j->sinks = sinks; j->ax = out_ax; j->ay = out_ay; j->tot = out_tot;
        j->lo = t * chunk;
        j->hi = (t + 1) * chunk < nsinks ? (t + 1) * chunk : nsinks;
        if (j->lo > j->hi) j->lo = j->hi;
    }
    if (nthreads == 1) {
        direct_worker(&jobs[0]);
    } else {
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, direct_worker, &jobs[t]);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* QuadTree restatement (quadtree.h:29-171, nBodySolverBH.h:33-116)                              */
/* ------------------------------------------------------------------------------------------ */
#define QT_MAX_DEPTH 64 /* the reference recurses without bound on >max_children coincident points */

typedef struct {
    double lowx, lowy, highx, highy;
    float total_mass;
    double comx, comy;
    int64_t occ_first, occ_count;
    int32_t child[4];
    int32_t nchild;
    int32_t depth;
    /* a node with exactly one child keeps its own occupant list (quadtree.h:140 never clears
     * it); descendants' partitions reorder the shared range, so the list is copied here */
    uint32_t* own_occ;
} qnode;

struct orc_quadtree {
    qnode* nodes;
    int64_t nn, cap;
    uint32_t* occ; /* occupant indices; each node's occupants are a contiguous range */
    uint32_t* tmp;
    const double *x, *y;
    int64_t n;
    int max_children;
};

static int32_t qt_new(orc_quadtree* t) {
    if (t->nn == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 1024;
        t->nodes = (qnode*)realloc(t->nodes, sizeof(qnode) * (size_t)t->cap);
    }
    qnode* q = &t->nodes[t->nn];
    memset(q, 0, sizeof(*q));
    return (int32_t)t->nn++;
}

/* createTree: returns the (float mass, float com) pair of quadtree.h:96 */
static void qt_create(orc_quadtree* t, int32_t id, float* ret_m, float* ret_cx, float* ret_cy) {
    qnode* q = &t->nodes[id];
    if (q->occ_count > t->max_children && q->depth < QT_MAX_DEPTH) {
        const double l = (q->highx - q->lowx) / 2.0;
        const double midx = q->lowx + l;
        const double midy = q->lowy + l;
        /* stable partition into HH, HL, LH, LL (quadtree.h:109-133) */
        int64_t cnt[4] = {0, 0, 0, 0};
        uint32_t* src = t->occ + q->occ_first;
        uint32_t* buf = t->tmp;
        int64_t pos = 0;
        for (int g = 0; g < 4; ++g) {
            for (int64_t i = 0; i < q->occ_count; ++i) {
                const uint32_t idx = src[i];
                const double px = t->x[idx], py = t->y[idx];
                int grp;
                if (px >= midx && py >= midy) grp = 0;
                else if (px >= midx && py < midy) grp = 1;
                else if (px < midx && py >= midy) grp = 2;
                else if (px < midx && py < midy) grp = 3;
                else grp = -1; /* NaN: the reference prints a message and drops the point */
                if (grp == g) {
                    buf[pos++] = idx;
                    cnt[g]++;
                }
            }
        }
        memcpy(src, buf, sizeof(uint32_t) * (size_t)pos);
        const double lowx = q->lowx, lowy = q->lowy, highx = q->highx, highy = q->highy;
        const int64_t base = q->occ_first;
        const int32_t depth = q->depth;
        int64_t off = 0;
        int32_t kids[4];
        int nk = 0;
        for (int g = 0; g < 4; ++g) {
            if (cnt[g] == 0) {
                continue;
            }
            int32_t c = qt_new(t);
            qnode* cq = &t->nodes[c];
            cq->occ_first = base + off;
            cq->occ_count = cnt[g];
            cq->depth = depth + 1;
            if (g == 0) { cq->lowx = midx; cq->lowy = midy; cq->highx = highx; cq->highy = highy; }
            if (g == 1) { cq->lowx = midx; cq->lowy = lowy; cq->highx = highx; cq->highy = midy; }
            if (g == 2) { cq->lowx = lowx; cq->lowy = midy; cq->highx = midx; cq->highy = highy; }
            if (g == 3) { cq->lowx = lowx; cq->lowy = lowy; cq->highx = midx; cq->highy = midy; }
            off += cnt[g];
            kids[nk++] = c;
        }
        q = &t->nodes[id];
        q->nchild = nk;
        for (int c = 0; c < nk; ++c) q->child[c] = kids[c];
        if (nk == 1) {
            q->own_occ = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)q->occ_count);
            memcpy(q->own_occ, t->occ + q->occ_first, sizeof(uint32_t) * (size_t)q->occ_count);
        }
        float tm = 0.0f;
        double cx = 0.0, cy = 0.0;
        for (int c = 0; c < nk; ++c) {
            float cm, ccx, ccy;
            qt_create(t, kids[c], &cm, &ccx, &ccy);
            tm = (float)((double)tm + (double)cm);
            cx += (double)cm * (double)ccx;
            cy += (double)cm * (double)ccy;
        }
        q = &t->nodes[id];
        q->total_mass = tm;
        q->comx = cx / (double)tm;
        q->comy = cy / (double)tm;
    } else {
        float tm = 0.0f;
        double cx = 0.0, cy = 0.0;
        for (int64_t i = 0; i < q->occ_count; ++i) {
            const uint32_t idx = t->occ[q->occ_first + i];
            tm = (float)((double)tm + 1.0);
            cx += 1.0 * t->x[idx];
            cy += 1.0 * t->y[idx];
        }
        q->total_mass = tm;
        q->comx = cx / (double)tm;
        q->comy = cy / (double)tm;
    }
    *ret_m = q->total_mass;
    *ret_cx = (float)q->comx;
    *ret_cy = (float)q->comy;
}

orc_quadtree* orc_quadtree_build(const double* x, const double* y, int64_t n, int max_children) {
    orc_quadtree* t = (orc_quadtree*)calloc(1, sizeof(orc_quadtree));
    t->x = x;
    t->y = y;
    t->n = n;
    t->max_children = max_children;
    t->occ = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    t->tmp = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    double lx = INFINITY, ly = INFINITY, hx = -INFINITY, hy = -INFINITY;
    for (int64_t i = 0; i < n; ++i) {
        t->occ[i] = (uint32_t)i;
        lx = gmin(lx, x[i]);
        ly = gmin(ly, y[i]);
        hx = gmax(hx, x[i]);
        hy = gmax(hy, y[i]);
    }
    const double largest = (hx - lx) > (hy - ly) ? (hx - lx) : (hy - ly); /* std::max */
    int32_t r = qt_new(t);
    qnode* q = &t->nodes[r];
    q->lowx = lx;
    q->lowy = ly;
    q->highx = lx + largest + 0.0001;
    q->highy = ly + largest + 0.0001;
    q->occ_first = 0;
    q->occ_count = n;
    q->depth = 0;
    float m, cx, cy;
    if (n > 0) qt_create(t, r, &m, &cx, &cy);
    return t;
}

void orc_quadtree_free(orc_quadtree* t) {
    if (!t) return;
    for (int64_t k = 0; k < t->nn; ++k) free(t->nodes[k].own_occ);
    free(t->nodes);
    free(t->occ);
    free(t->tmp);
    free(t);
}

int64_t orc_quadtree_node_count(const orc_quadtree* t) { return t->nn; }

/* pre-order walk (children in HH, HL, LH, LL order) */
static void qt_export_rec(const orc_quadtree* t, int32_t id, int64_t* k, int32_t* depth,
                          uint32_t* count, uint32_t* nchild, double* lowx, double* lowy,
                          double* highx, double* highy, double* mass, double* comx, double* comy,
                          uint32_t* node_first) {
    const qnode* q = &t->nodes[id];
    const int64_t me = (*k)++;
    if (depth) depth[me] = q->depth;
    if (count) count[me] = (uint32_t)q->occ_count;
    if (nchild) nchild[me] = (uint32_t)q->nchild;
    if (lowx) lowx[me] = q->lowx;
    if (lowy) lowy[me] = q->lowy;
    if (highx) highx[me] = q->highx;
    if (highy) highy[me] = q->highy;
    if (mass) mass[me] = (double)q->total_mass;
    if (comx) comx[me] = q->comx;
    if (comy) comy[me] = q->comy;
    if (node_first) node_first[me] = (uint32_t)q->occ_first;
    for (int c = 0; c < q->nchild; ++c)
        qt_export_rec(t, q->child[c], k, depth, count, nchild, lowx, lowy, highx, highy, mass,
                      comx, comy, node_first);
}

void orc_quadtree_export(const orc_quadtree* t, int32_t* depth, uint32_t* count, uint32_t* nchild,
                         double* lowx, double* lowy, double* highx, double* highy, double* mass,
                         double* comx, double* comy) {
    int64_t k = 0;
    if (t->nn) qt_export_rec(t, 0, &k, depth, count, nchild, lowx, lowy, highx, highy, mass, comx,
                             comy, NULL);
}

void orc_quadtree_occupants(const orc_quadtree* t, uint32_t* node_first, uint32_t* occ_index) {
    int64_t k = 0;
    if (t->nn) qt_export_rec(t, 0, &k, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL,
                             node_first);
    memcpy(occ_index, t->occ, sizeof(uint32_t) * (size_t)t->n);
}

typedef struct {
    const orc_quadtree* t;
    int kernel;
    double G, m, eps2, theta;
    double *ax, *ay;
    int64_t lo, hi;
    double total;
    uint64_t pn, ppc;
} qt_job;

/* traverseBH (nBodySolverBH.h:56-89) */
static void qt_traverse(qt_job* j, int64_t i, int32_t id, double* ax, double* ay, double* tot) {
    const orc_quadtree* t = j->t;
    const qnode* q = &t->nodes[id];
    const double px = t->x[i], py = t->y[i];
    const double l = q->highx - q->lowx;
    const double dx = px - q->comx, dy = py - q->comy;
    const double len = sqrt(dx * dx + dy * dy);
    if (l / len < j->theta) {
        const double M = (double)q->total_mass * j->m;
        interact(j->kernel, j->G, j->eps2, px, py, q->comx, q->comy, M, ax, ay, tot);
        j->pn++;
    } else if (q->nchild <= 1) {
        const uint32_t* list = q->own_occ ? q->own_occ : t->occ + q->occ_first;
        for (int64_t k = 0; k < q->occ_count; ++k) {
            const uint32_t o = list[k];
            if ((int64_t)o == i) continue;
            interact(j->kernel, j->G, j->eps2, px, py, t->x[o], t->y[o], j->m, ax, ay, tot);
            j->ppc++;
        }
    } else {
        for (int c = 0; c < q->nchild; ++c) qt_traverse(j, i, q->child[c], ax, ay, tot);
    }
}

static void* qt_worker(void* arg) {
    qt_job* j = (qt_job*)arg;
    double total = 0.0;
    for (int64_t i = j->lo; i < j->hi; ++i) {
        double ax = j->ax[i], ay = j->ay[i];
        qt_traverse(j, i, 0, &ax, &ay, &total);
        j->ax[i] = ax;
        j->ay[i] = ay;
    }
    j->total = total;
    return NULL;
}

double orc_quadtree_forces(const orc_quadtree* t, const double* x, const double* y, int kernel,
                           double G, double m, double eps2, double theta, double* out_ax,
                           double* out_ay, uint64_t* counters, int nthreads) {
    (void)x;
    (void)y;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    qt_job jobs[256];
    pthread_t th[256];
    const int64_t n = t->n;
    const int64_t chunk = (n + nthreads - 1) / nthreads;
    for (int k = 0; k < nthreads; ++k) {
        qt_job* j = &jobs[k];
        j->t = t; j->kernel = kernel; j->G = G; j->m = m; j->eps2 = eps2; j->theta = theta;
        j->ax = out_ax; j->ay = out_ay;
        j->lo = k * chunk;
        j->hi = (k + 1) * chunk < n ? (k + 1) * chunk : n;
        if (j->lo > j->hi) j->lo = j->hi;
        j->total = 0; j->pn = j->ppc = 0;
    }
    if (t->nn == 0) return 0.0;
    if (nthreads == 1) {
        qt_worker(&jobs[0]);
    } else {
        for (int k = 0; k < nthreads; ++k) pthread_create(&th[k], NULL, qt_worker, &jobs[k]);
        for (int k = 0; k < nthreads; ++k) pthread_join(th[k], NULL);
    }
    double total = 0.0;
    if (counters) counters[0] = counters[1] = 0;
    for (int k = 0; k < nthreads; ++k) {
        total += jobs[k].total;
        if (counters) {
            counters[0] += jobs[k].pn;
            counters[1] += jobs[k].ppc;
        }
    }
    return total;
}

/* ------------------------------------------------------------------------------------------ */
/* leapfrog with BH forces                                                                       */
/* ------------------------------------------------------------------------------------------ */
/* QuadTree root side: ((min + largest) + 1e-4) - min (quadtree.h:45-47, nBodySolverBH.h:58) */
static double qt_root_side(double minx, double miny, double maxx, double maxy) {
    const double largest = (maxx - minx) > (maxy - miny) ? (maxx - minx) : (maxy - miny);
    return ((minx + largest) + 0.0001) - minx;
}

static void bh_forces_pipeline(const double* x, const double* y, const double* m, int64_t n,
                               double G, double eps2, double theta, int semantics, int leaf_cap,
                               int bits, int nthreads, double* ax, double* ay) {
    double minx = INFINITY, miny = INFINITY, maxx = -INFINITY, maxy = -INFINITY;
    for (int64_t i = 0; i < n; ++i) {
        minx = gmin(minx, x[i]);
        miny = gmin(miny, y[i]);
        maxx = gmax(maxx, x[i]);
        maxy = gmax(maxy, y[i]);
    }
    uint64_t* code = (uint64_t*)malloc(sizeof(uint64_t) * n);
    uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * n);
    double side;
    int depth = 0;
    if (semantics == ORC_SEM_QUADTREE) {
        side = qt_root_side(minx, miny, maxx, maxy);
        orc_morton64(x, y, n, minx, miny, side, bits, code);
    } else {
        side = gmax(maxx - minx, maxy - miny);
        uint32_t* c32 = (uint32_t*)malloc(sizeof(uint32_t) * n);
        orc_morton32(x, y, n, minx, miny, side, c32);
        for (int64_t i = 0; i < n; ++i) code[i] = c32[i];
        free(c32);
        depth = orc_get_depth(leaf_cap, n);
    }
    orc_stable_sort_u64(code, n, perm);
    double* xs = (double*)malloc(sizeof(double) * n);
    double* ys = (double*)malloc(sizeof(double) * n);
    double* ms = (double*)malloc(sizeof(double) * n);
    uint64_t* cs = (uint64_t*)malloc(sizeof(uint64_t) * n);
    uint32_t* cs32 = (uint32_t*)malloc(sizeof(uint32_t) * n);
    for (int64_t i = 0; i < n; ++i) {
        xs[i] = x[perm[i]];
        ys[i] = y[perm[i]];
        ms[i] = m ? m[perm[i]] : 1.0;
        cs[i] = code[perm[i]];
        cs32[i] = (uint32_t)cs[i];
    }
    const int64_t cap = 3 * n + 64 + orc_fixed_tree_slots(depth);
    int32_t* level = (int32_t*)malloc(sizeof(int32_t) * cap);
    uint32_t* first = (uint32_t*)malloc(sizeof(uint32_t) * cap);
    uint32_t* count = (uint32_t*)malloc(sizeof(uint32_t) * cap);
    uint32_t* skip = (uint32_t*)malloc(sizeof(uint32_t) * cap);
    uint32_t* nchild = (uint32_t*)malloc(sizeof(uint32_t) * cap);
    int64_t nn;
    if (semantics == ORC_SEM_QUADTREE)
        nn = orc_adaptive_tree(cs, n, bits, leaf_cap, cap, level, first, count, skip, nchild);
    else
        nn = orc_fixed_preorder(cs32, n, depth, cap, level, first, count, skip, nchild);
    uint8_t* pp = (uint8_t*)malloc(nn > 0 ? nn : 1);
    double* cx = (double*)malloc(sizeof(double) * (nn > 0 ? nn : 1));
    double* cy = (double*)malloc(sizeof(double) * (nn > 0 ? nn : 1));
    double* mass = (double*)malloc(sizeof(double) * (nn > 0 ? nn : 1));
    for (int64_t k = 0; k < nn; ++k)
        pp[k] = semantics == ORC_SEM_QUADTREE ? (nchild[k] <= 1) : (level[k] == depth);
    orc_node_moments(xs, ys, ms, nn, first, count, cx, cy, mass, NULL);
    double* sax = (double*)malloc(sizeof(double) * n);
    double* say = (double*)malloc(sizeof(double) * n);
    orc_bh_traverse(xs, ys, ms, n, level, first, count, skip, pp, cx, cy, mass, nn, side,
                    semantics, ORC_KERNEL_GRAVITY, G, eps2, theta, NULL, n, sax, say, NULL, NULL,
                    nthreads);
    for (int64_t i = 0; i < n; ++i) {
        ax[perm[i]] = sax[i];
        ay[perm[i]] = say[i];
    }
    free(code); free(perm); free(xs); free(ys); free(ms); free(cs); free(cs32);
    free(level); free(first); free(count); free(skip); free(nchild);
    free(pp); free(cx); free(cy); free(mass); free(sax); free(say);
}

/* kick-drift-kick; the reference's dead gravity integrator is kick-then-drift
 * (gravitysim.h:99-121); KDK is the time-symmetric form used here. */
void orc_leapfrog(double* x, double* y, double* vx, double* vy, const double* m, int64_t n,
                  int steps, double dt, double G, double eps2, double theta, int semantics,
                  int leaf_cap, int bits, int nthreads) {
    double* ax = (double*)malloc(sizeof(double) * n);
    double* ay = (double*)malloc(sizeof(double) * n);
    bh_forces_pipeline(x, y, m, n, G, eps2, theta, semantics, leaf_cap, bits, nthreads, ax, ay);
    for (int s = 0; s < steps; ++s) {
        for (int64_t i = 0; i < n; ++i) {
            vx[i] += 0.5 * dt * ax[i];
            vy[i] += 0.5 * dt * ay[i];
            x[i] += dt * vx[i];
            y[i] += dt * vy[i];
        }
        bh_forces_pipeline(x, y, m, n, G, eps2, theta, semantics, leaf_cap, bits, nthreads, ax,
                           ay);
        for (int64_t i = 0; i < n; ++i) {
            vx[i] += 0.5 * dt * ax[i];
            vy[i] += 0.5 * dt * ay[i];
        }
    }
    free(ax);
    free(ay);
}

/* ------------------------------------------------------------------------------------------ */
/* t-SNE                                                                                         */
/* ------------------------------------------------------------------------------------------ */

/* thetaFunction (tsne.h:208-233) */
double orc_tsne_theta(double min_theta, double max_theta, int iteration) {
    const double sharpness = 1.0;
    double r = pow((double)iteration / 999.0, sharpness);
    if (r > 1.0) r = 1.0;
    const double extra = 0.5 + 0.5 * cos(M_PI * r);
    return min_theta + extra * (max_theta - min_theta);
}

/* updateAttractive (tsne.h:356-381) */
void orc_tsne_attraction(const double* x, const double* y, int64_t n, const uint32_t* row_ptr,
                         const uint32_t* col, const double* val, double exag, double* dx,
                         double* dy) {
    for (int64_t r = 0; r < n; ++r) {
        double sx = 0.0, sy = 0.0;
        for (uint32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
            const uint32_t c = col[e];
            const double ddx = x[c] - x[r], ddy = y[c] - y[r];
            const double d = ddx * ddx + ddy * ddy;
            const double q = 1.0 / (1.0 + d);
            const double s = exag * val[e] * q;
            sx += s * ddx;
            sy += s * ddy;
        }
        dx[r] += sx;
        dy[r] += sy;
    }
}

typedef struct {
    const double *x, *y;
    int64_t n, lo, hi;
    double z;
} z_job;

static void* z_worker(void* arg) {
    z_job* j = (z_job*)arg;
    double s = 0.0, c = 0.0;
    for (int64_t i = j->lo; i < j->hi; ++i) {
        for (int64_t q = 0; q < j->n; ++q) {
            if (q == i) continue;
            const double dx = j->x[q] - j->x[i], dy = j->y[q] - j->y[i];
            kahan_add(&s, &c, 1.0 / (1.0 + (dx * dx + dy * dy)));
        }
    }
    j->z = s;
    return NULL;
}

double orc_tsne_z(const double* x, const double* y, int64_t n, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    z_job jobs[256];
    pthread_t th[256];
    const int64_t chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].x = x; jobs[t].y = y; jobs[t].n = n;
        jobs[t].lo = t * chunk;
        jobs[t].hi = (t + 1) * chunk < n ? (t + 1) * chunk : n;
        if (jobs[t].lo > jobs[t].hi) jobs[t].lo = jobs[t].hi;
        jobs[t].z = 0;
        pthread_create(&th[t], NULL, z_worker, &jobs[t]);
    }
    double z = 0.0;
    for (int t = 0; t < nthreads; ++t) {
        pthread_join(th[t], NULL);
        z += jobs[t].z;
    }
    return z;
}

/* costFunction (tsne.h:383-460; tsneTests.h:1388-1456): log base 2 */
double orc_tsne_kl(const double* x, const double* y, int64_t n, const uint32_t* row_ptr,
                   const uint32_t* col, const double* val, int nthreads) {
    const double z = orc_tsne_z(x, y, n, nthreads);
    double cost = 0.0;
    for (int64_t r = 0; r < n; ++r) {
        for (uint32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
            const uint32_t c = col[e];
            if ((int64_t)c == r) continue;
            const double dx = x[c] - x[r], dy = y[c] - y[r];
            const double q = (1.0 / (1.0 + (dx * dx + dy * dy))) / z;
            const double p = val[e];
            cost += p * log2(p / q);
        }
    }
    return cost;
}

/* One iteration at iteration counter t: updateDerivative (tsne.h:272-301, 356-381) with the
 * reference BH solver (QuadTree, leaf capacity 16, t-SNE kernels, nBodySolverBH.h:33-116), then
 * updatePoints (tsne.h:179-206): y_next = y + lr d + mu (y - y_prev). Writes d (may be NULL)
 * and y_next; returns the Barnes-Hut Z. */
double orc_tsne_step(const double* x, const double* y, const double* px, const double* py,
                     int64_t n, const uint32_t* row_ptr, const uint32_t* col, const double* val,
                     double theta, double exag, double lr, double mu, double* dx, double* dy,
                     double* nx, double* ny, int nthreads) {
    double* rx = (double*)calloc((size_t)n, sizeof(double));
    double* ry = (double*)calloc((size_t)n, sizeof(double));
    orc_quadtree* qt = orc_quadtree_build(x, y, n, 16);
    const double z = orc_quadtree_forces(qt, x, y, 1, 1.0, 1.0, 0.0, theta, rx, ry, NULL,
                                         nthreads);
    orc_quadtree_free(qt);
    const double s = 4.0 / z;
    for (int64_t i = 0; i < n; ++i) {
        rx[i] *= s;
        ry[i] *= s;
    }
    orc_tsne_attraction(x, y, n, row_ptr, col, val, exag, rx, ry);
    for (int64_t i = 0; i < n; ++i) {
        if (dx) dx[i] = rx[i];
        if (dy) dy[i] = ry[i];
        nx[i] = x[i] + lr * rx[i] + mu * (x[i] - px[i]);
        ny[i] = y[i] + lr * ry[i] + mu * (y[i] - py[i]);
    }
    free(rx);
    free(ry);
    return z;
}

/* ------------------------------------------------------------------------------------------ */
/* MLP (din -> h -> h -> dout, SiLU), FP64 reference for the learned far-field correction       */
/* ------------------------------------------------------------------------------------------ */
static double silu(double v) { return v / (1.0 + exp(-v)); }
static double dsilu(double v) {
    const double s = 1.0 / (1.0 + exp(-v));
    return s * (1.0 + v * (1.0 - s));
}

static void mlp_fwd_one(const double* xin, int din, int h, int dout, const double* W1,
                        const double* b1, const double* W2, const double* b2, const double* W3,
                        const double* b3, double* z1, double* a1, double* z2, double* a2,
                        double* y) {
    for (int j = 0; j < h; ++j) {
        double s = b1[j];
        for (int i = 0; i < din; ++i) s += xin[i] * W1[i * h + j];
        z1[j] = s;
        a1[j] = silu(s);
    }
    for (int j = 0; j < h; ++j) {
        double s = b2[j];
        for (int i = 0; i < h; ++i) s += a1[i] * W2[i * h + j];
        z2[j] = s;
        a2[j] = silu(s);
    }
    for (int o = 0; o < dout; ++o) {
        double s = b3[o];
        for (int i = 0; i < h; ++i) s += a2[i] * W3[i * dout + o];
        y[o] = s;
    }
}

void orc_mlp_forward(const double* X, int64_t B, int din, int h, int dout, const double* W1,
                     const double* b1, const double* W2, const double* b2, const double* W3,
                     const double* b3, double* Y) {
    double* buf = (double*)malloc(sizeof(double) * 4 * h);
    for (int64_t b = 0; b < B; ++b)
        mlp_fwd_one(X + b * din, din, h, dout, W1, b1, W2, b2, W3, b3, buf, buf + h, buf + 2 * h,
                    buf + 3 * h, Y + b * dout);
    free(buf);
}

/* backward of one sample with output gradient dy, accumulated into the parameter gradients */
static void mlp_bwd_one(const double* xin, int din, int h, int dout, const double* W2,
                        const double* W3, const double* z1, const double* a1, const double* z2,
                        const double* a2, const double* dy, double* d1, double* d2, double* gW1,
                        double* gb1, double* gW2, double* gb2, double* gW3, double* gb3) {
    for (int o = 0; o < dout; ++o) gb3[o] += dy[o];
    for (int i = 0; i < h; ++i) {
        double s = 0.0;
        for (int o = 0; o < dout; ++o) {
            gW3[i * dout + o] += a2[i] * dy[o];
            s += W3[i * dout + o] * dy[o];
        }
        d2[i] = s * dsilu(z2[i]);
        gb2[i] += d2[i];
    }
    for (int i = 0; i < h; ++i) {
        double s = 0.0;
        for (int j = 0; j < h; ++j) {
            gW2[i * h + j] += a1[i] * d2[j];
            s += W2[i * h + j] * d2[j];
        }
        d1[i] = s * dsilu(z1[i]);
        gb1[i] += d1[i];
    }
    for (int i = 0; i < din; ++i)
        for (int j = 0; j < h; ++j) gW1[i * h + j] += xin[i] * d1[j];
}

double orc_mlp_loss_grad(const double* X, const double* T, int64_t B, int din, int h, int dout,
                         const double* W1, const double* b1, const double* W2, const double* b2,
                         const double* W3, const double* b3, double* gW1, double* gb1,
                         double* gW2, double* gb2, double* gW3, double* gb3) {
    memset(gW1, 0, sizeof(double) * din * h);
    memset(gb1, 0, sizeof(double) * h);
    memset(gW2, 0, sizeof(double) * h * h);
    memset(gb2, 0, sizeof(double) * h);
    memset(gW3, 0, sizeof(double) * h * dout);
    memset(gb3, 0, sizeof(double) * dout);
    double* buf = (double*)malloc(sizeof(double) * (6 * h + 2 * dout));
    double *z1 = buf, *a1 = buf + h, *z2 = buf + 2 * h, *a2 = buf + 3 * h, *d2 = buf + 4 * h,
           *d1 = buf + 5 * h, *y = buf + 6 * h, *dy = buf + 6 * h + dout;
    const double scale = 1.0 / ((double)B * (double)dout);
    double loss = 0.0;
    for (int64_t b = 0; b < B; ++b) {
        const double* xin = X + b * din;
        mlp_fwd_one(xin, din, h, dout, W1, b1, W2, b2, W3, b3, z1, a1, z2, a2, y);
        for (int o = 0; o < dout; ++o) {
            const double e = y[o] - T[b * dout + o];
            loss += e * e * scale;
            dy[o] = 2.0 * e * scale;
            gb3[o] += dy[o];
        }
        for (int i = 0; i < h; ++i) {
            double s = 0.0;
            for (int o = 0; o < dout; ++o) {
                gW3[i * dout + o] += a2[i] * dy[o];
                s += W3[i * dout + o] * dy[o];
            }
            d2[i] = s * dsilu(z2[i]);
            gb2[i] += d2[i];
        }
        for (int i = 0; i < h; ++i) {
            double s = 0.0;
            for (int j = 0; j < h; ++j) {
                gW2[i * h + j] += a1[i] * d2[j];
                s += W2[i * h + j] * d2[j];
            }
            d1[i] = s * dsilu(z1[i]);
            gb1[i] += d1[i];
        }
        for (int i = 0; i < din; ++i)
            for (int j = 0; j < h; ++j) gW1[i * h + j] += xin[i] * d1[j];
    }
    free(buf);
    return loss;
}

/* Adam (Kingma & Ba 2015, Algorithm 1) */
void orc_adam(double* p, const double* g, double* mom1, double* mom2, int64_t n, double lr,
              double beta1, double beta2, double eps, int64_t t) {
    const double c1 = 1.0 - pow(beta1, (double)t);
    const double c2 = 1.0 - pow(beta2, (double)t);
    for (int64_t i = 0; i < n; ++i) {
        mom1[i] = beta1 * mom1[i] + (1.0 - beta1) * g[i];
        mom2[i] = beta2 * mom2[i] + (1.0 - beta2) * g[i] * g[i];
        const double mh = mom1[i] / c1;
        const double vh = mom2[i] / c2;
        p[i] -= lr * mh / (sqrt(vh) + eps);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Symmetric dual-tree FMM on the fixed-depth Morton tree, t-SNE kernels (NBodySolverFMM_SYM_   */
/* MORTON: traverse_SYM_NN :154-270, traverse_SYM_PN :272-313, applyForces :482-536, solveNbody */
/* :68-76; TSNE_FMM_SYM_MORTON_NN/PN/PP_Kernel :622-836), in FP64 over the slots of            */
/* orc_fixed_tree. Expansions per node: C1 (2), C2 (2x2, [i][j] at 2i+j), C3 (2x2x2, [i][j][k] */
/* at 4i+2j+k).                                                                                 */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    const double *xs, *ys;
    const uint32_t *first_child, *first, *count;
    const double *bb_len, *cx, *cy, *m0, *m2;
    double theta;
    double *c1, *c2, *c3; /* 2, 4, 8 per slot */
    double *ax, *ay;
    double total;
} fmm_ctx;

static void fmm_derivs(double rx, double ry, double* rS, double* D1, double* D2, double* D3) {
    const double sq = rx * rx + ry * ry;
    *rS = 1.0 + sq;
    const double inv = 1.0 / *rS, inv2 = inv * inv, inv3 = inv2 * inv, inv4 = inv3 * inv;
    *D1 = inv2;
    *D2 = -4.0 * inv3;
    *D3 = 24.0 * inv4;
}

/* C1 of a sink at offset R from a source of mass mb and second moment m2 (xx, xy, yy) */
static void fmm_c1(double rx, double ry, double D1, double D2, double D3, double mb,
                   const double* m2, double* out) {
    const double t00 = m2[0] / mb, t01 = m2[1] / mb, t10 = m2[1] / mb, t11 = m2[2] / mb;
    const double s1 = t00 + t11;
    const double s2 = (rx * rx * t00) + (rx * ry * t01) + (ry * rx * t10) + (ry * ry * t11);
    const double s3x = rx * t00 + ry * t01, s3y = rx * t10 + ry * t11;
    out[0] = mb * (rx * (D1 + 0.5 * s1 * D2 + 0.5 * s2 * D3) + s3x * D2);
    out[1] = mb * (ry * (D1 + 0.5 * s1 * D2 + 0.5 * s2 * D3) + s3y * D2);
}

static void fmm_c2(double rx, double ry, double D1, double D2, double f, double* c2) {
    c2[0] += f * (D1 + rx * rx * D2);
    c2[1] += f * (rx * ry * D2);
    c2[2] += f * (ry * rx * D2);
    c2[3] += f * (D1 + ry * ry * D2);
}

static void fmm_c3(double rx, double ry, double D2, double D3, double f, double* c3) {
    c3[0] += f * ((rx + rx + rx) * D2 + rx * rx * rx * D3); /* 000 */
    c3[1] += f * ((ry)*D2 + rx * rx * ry * D3);            /* 001 */
    c3[2] += f * ((ry)*D2 + rx * ry * rx * D3);            /* 010 */
    c3[3] += f * ((rx)*D2 + rx * ry * ry * D3);            /* 011 */
    c3[4] += f * ((ry)*D2 + ry * rx * rx * D3);            /* 100 */
    c3[5] += f * ((rx)*D2 + ry * rx * ry * D3);            /* 101 */
    c3[6] += f * ((rx)*D2 + ry * ry * rx * D3);            /* 110 */
    c3[7] += f * ((ry + ry + ry) * D2 + ry * ry * ry * D3); /* 111 */
}

/* TSNE_FMM_SYM_MORTON_NN_Kernel (:622-723) */
static void fmm_nn(fmm_ctx* c, int64_t a, int64_t b) {
    double rx = c->cx[a] - c->cx[b], ry = c->cy[a] - c->cy[b];
    double rS, D1, D2, D3;
    fmm_derivs(rx, ry, &rS, &D1, &D2, &D3);
    double v[2];
    fmm_c1(rx, ry, D1, D2, D3, c->m0[b], c->m2 + 3 * b, v);
    c->c1[2 * a] += v[0];
    c->c1[2 * a + 1] += v[1];
    double c2[4] = {0, 0, 0, 0}, c3[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    fmm_c2(rx, ry, D1, D2, c->m0[a] * c->m0[b], c2);
    fmm_c3(rx, ry, D2, D3, c->m0[a] * c->m0[b], c3);
    for (int k = 0; k < 4; ++k) c->c2[4 * a + k] += c2[k];
    for (int k = 0; k < 8; ++k) c->c3[8 * a + k] += c3[k];
    /* second interaction: R -> -R, the other node's moments; C2 added as is, C3 negated */
    fmm_c1(-rx, -ry, D1, D2, D3, c->m0[a], c->m2 + 3 * a, v);
    c->c1[2 * b] += v[0];
    c->c1[2 * b + 1] += v[1];
    for (int k = 0; k < 4; ++k) c->c2[4 * b + k] += c2[k];
    for (int k = 0; k < 8; ++k) c->c3[8 * b + k] += -c3[k];
    c->total += 2.0 * (c->m0[a] * c->m0[b]) / rS;
}

/* TSNE_FMM_SYM_MORTON_PN_Kernel (:726-816) */
static void fmm_pn(fmm_ctx* c, int64_t p, int64_t b) {
    double rx = c->xs[p] - c->cx[b], ry = c->ys[p] - c->cy[b];
    double rS, D1, D2, D3;
    fmm_derivs(rx, ry, &rS, &D1, &D2, &D3);
    double v[2];
    fmm_c1(rx, ry, D1, D2, D3, c->m0[b], c->m2 + 3 * b, v);
    c->ax[p] += v[0];
    c->ay[p] += v[1];
    rx = -rx;
    ry = -ry;
    c->c1[2 * b] += rx * D1;
    c->c1[2 * b + 1] += ry * D1;
    fmm_c2(rx, ry, D1, D2, c->m0[b], c->c2 + 4 * b);
    fmm_c3(rx, ry, D2, D3, c->m0[b], c->c3 + 8 * b);
    c->total += 2.0 * c->m0[b] / rS;
}

/* TSNE_FMM_SYM_MORTON_PP_Kernel (:818-836) */
static void fmm_pp(fmm_ctx* c, int64_t p, int64_t q) {
    const double rx = c->xs[p] - c->xs[q], ry = c->ys[p] - c->ys[q];
    const double w = 1.0 / (1.0 + (rx * rx + ry * ry));
    c->ax[p] += w * w * rx;
    c->ay[p] += w * w * ry;
    c->ax[q] += w * w * -rx;
    c->ay[q] += w * w * -ry;
    c->total += 2.0 * w;
}

static void fmm_sym_pn(fmm_ctx* c, int64_t p, int64_t b) {
    const double rx = c->xs[p] - c->cx[b], ry = c->ys[p] - c->cy[b];
    const double dist = rx * rx + ry * ry;
    const double thr = c->bb_len[b] / c->theta;
    if (dist > thr * thr) {
        fmm_pn(c, p, b);
    } else if (c->first_child[b] == 0) {
        for (int64_t q = c->first[b]; q < (int64_t)c->first[b] + c->count[b]; ++q)
            if (q != p) fmm_pp(c, p, q);
    } else {
        for (int64_t ch = c->first_child[b]; ch < (int64_t)c->first_child[b] + 4; ++ch)
            if (c->count[ch] != 0) fmm_sym_pn(c, p, ch);
    }
}

static void fmm_sym_nn(fmm_ctx* c, int64_t a, int64_t b) {
    const double rx = c->cx[a] - c->cx[b], ry = c->cy[a] - c->cy[b];
    const double dist = rx * rx + ry * ry;
    const double thr = (c->bb_len[a] + c->bb_len[b]) / c->theta;
    if (dist > thr * thr) {
        if (c->count[a] != 0 && c->count[b] != 0) fmm_nn(c, a, b);
    } else if (c->first_child[a] == 0) {
        if (a == b) {
            for (int64_t p = c->first[a]; p < (int64_t)c->first[a] + c->count[a]; ++p)
                if (c->count[b] != 0)
                    for (int64_t q = p; q < (int64_t)c->first[b] + c->count[b]; ++q)
                        if (p != q) fmm_pp(c, p, q);
        } else {
            for (int64_t p = c->first[a]; p < (int64_t)c->first[a] + c->count[a]; ++p)
                if (c->count[b] != 0) fmm_sym_pn(c, p, b);
        }
    } else if (c->first_child[b] == 0) {
        if (a == b) {
            for (int64_t q = c->first[b]; q < (int64_t)c->first[b] + c->count[b]; ++q)
                if (c->count[a] != 0)
                    for (int64_t p = q; p < (int64_t)c->first[a] + c->count[a]; ++p)
                        if (p != q) fmm_pp(c, p, q);
        } else {
            for (int64_t q = c->first[b]; q < (int64_t)c->first[b] + c->count[b]; ++q)
                if (c->count[a] != 0) fmm_sym_pn(c, q, a);
        }
    } else if (a == b) {
        for (int64_t i = c->first_child[a]; i < (int64_t)c->first_child[a] + 4; ++i)
            for (int64_t j = i; j < (int64_t)c->first_child[a] + 4; ++j)
                if (c->count[i] != 0 && c->count[j] != 0) fmm_sym_nn(c, i, j);
    } else {
        for (int64_t i = c->first_child[a]; i < (int64_t)c->first_child[a] + 4; ++i)
            for (int64_t j = c->first_child[b]; j < (int64_t)c->first_child[b] + 4; ++j)
                if (c->count[i] != 0 && c->count[j] != 0) fmm_sym_nn(c, i, j);
    }
}

/* v = C1 + d.C2 + 0.5 d.(d.C3) (contractions over the first index, :497-504) */
static void fmm_translate(const double* c1, const double* c2, const double* c3, double dx,
                          double dy, double* v1, double* v2) {
    const double d[2] = {dx, dy};
    double t[4]; /* t[j][k] = sum_i d_i C3[i][j][k] */
    for (int j = 0; j < 2; ++j)
        for (int k = 0; k < 2; ++k) t[2 * j + k] = d[0] * c3[0 * 4 + 2 * j + k] + d[1] * c3[1 * 4 + 2 * j + k];
    for (int k = 0; k < 2; ++k)
        v1[k] = c1[k] + (d[0] * c2[0 * 2 + k] + d[1] * c2[1 * 2 + k]) +
                0.5 * (d[0] * t[0 * 2 + k] + d[1] * t[1 * 2 + k]);
    if (v2)
        for (int k = 0; k < 4; ++k) v2[k] = c2[k] + t[k];
}

static void fmm_apply(fmm_ctx* c, int64_t node) {
    if (c->first_child[node] != 0) {
        for (int64_t ch = c->first_child[node]; ch < (int64_t)c->first_child[node] + 4; ++ch) {
            if (c->count[ch] == 0) continue;
            const double inv = 1.0 / c->m0[ch];
            for (int k = 0; k < 4; ++k) c->c2[4 * ch + k] *= inv;
            for (int k = 0; k < 8; ++k) c->c3[8 * ch + k] *= inv;
            double v1[2], v2[4];
            fmm_translate(c->c1 + 2 * node, c->c2 + 4 * node, c->c3 + 8 * node,
                          c->cx[ch] - c->cx[node], c->cy[ch] - c->cy[node], v1, v2);
            c->c1[2 * ch] += v1[0];
            c->c1[2 * ch + 1] += v1[1];
            for (int k = 0; k < 4; ++k) c->c2[4 * ch + k] += v2[k];
            for (int k = 0; k < 8; ++k) c->c3[8 * ch + k] += c->c3[8 * node + k];
            fmm_apply(c, ch);
        }
    } else {
        for (int64_t p = c->first[node]; p < (int64_t)c->first[node] + c->count[node]; ++p) {
            double v1[2];
            fmm_translate(c->c1 + 2 * node, c->c2 + 4 * node, c->c3 + 8 * node,
                          c->xs[p] - c->cx[node], c->ys[p] - c->cy[node], v1, NULL);
            c->ax[p] += v1[0];
            c->ay[p] += v1[1];
        }
    }
}

/* solveNbody (:68-76): repulsive derivative per sorted point (accumulated into ax, ay, which
 * the caller zeroes) and the total Z. Tree arrays as written by orc_fixed_tree. */
double orc_fmm_sym(const double* xs, const double* ys, int64_t n, int depth,
                   const uint32_t* first_child, const uint32_t* first_particle,
                   const uint32_t* count, const double* bb_len, const double* com_x,
                   const double* com_y, const double* m0, const double* m2, double theta,
                   double* ax, double* ay) {
    (void)n;
    const int64_t slots = orc_fixed_tree_slots(depth);
    fmm_ctx c;
    c.xs = xs;
    c.ys = ys;
    c.first_child = first_child;
    c.first = first_particle;
    c.count = count;
    c.bb_len = bb_len;
    c.cx = com_x;
    c.cy = com_y;
    c.m0 = m0;
    c.m2 = m2;
    c.theta = theta;
    c.c1 = (double*)calloc((size_t)slots * 2, sizeof(double));
    c.c2 = (double*)calloc((size_t)slots * 4, sizeof(double));
    c.c3 = (double*)calloc((size_t)slots * 8, sizeof(double));
    c.ax = ax;
    c.ay = ay;
    c.total = 0.0;
    if (count[0] != 0) {
        fmm_sym_nn(&c, 0, 0);
        const double inv = 1.0 / m0[0];
        for (int k = 0; k < 4; ++k) c.c2[k] *= inv;
        for (int k = 0; k < 8; ++k) c.c3[k] *= inv;
        fmm_apply(&c, 0);
    }
    free(c.c1);
    free(c.c2);
    free(c.c3);
    return c.total;
}

/* ------------------------------------------------------------------------------------------ */
/* learned far-field correction (no reference counterpart)                                      */
/* ------------------------------------------------------------------------------------------ */
/* Per cell: mass M, centre of mass c, side l, count n; s_j = x_j - c over the cell's particles.
 * Features (ORC_CELL_FEATURES):
 *   0..2  S/(M l^2):  sum m s_x s_x, sum m s_x s_y, sum m s_y s_y
 *   3..6  T/(M l^3):  sum m s_x^3, sum m s_x^2 s_y, sum m s_x s_y^2, sum m s_y^3
 *   7, 8  (c - box centre) / l, the box of the cell's level-L Morton prefix
 *   9     log2(n) / 16
 * Pair (sink p, accepted cell): d = c - p, r = |d|, w = d / r (complex), q = l / r,
 * scale s = M l^2 / r^4. The correction to the monopole (before G) is s * sum_j c_j b_j with the
 * network's 7 complex outputs c_j and the basis b = (w, conj w, w^3, q, q w^2, q conj(w)^2,
 * q w^4): the angular forms of the quadrupole (q^0) and octupole (q^1) terms of the Taylor
 * expansion of (d + s_j)/|d + s_j|^3 about s = 0. With c_1 = 0.75 tr(S), c_2 = 0.375 Q,
 * c_3 = 1.875 conj(Q), Q = S_xx - S_yy + 2i S_xy (S normalised as in the features), the
 * q^0 part is exactly the quadrupole term. */
static void deinterleave_prefix(uint64_t code, int bits, int level, uint64_t* ix, uint64_t* iy) {
    const uint64_t pre = level > 0 ? code >> (2 * (bits - level)) : 0;
    uint64_t x = 0, y = 0;
    for (int b = 0; b < level; ++b) {
        x |= ((pre >> (2 * b)) & 1ull) << b;
        y |= ((pre >> (2 * b + 1)) & 1ull) << b;
    }
    *ix = x;
    *iy = y;
}

void orc_cell_features(const double* xs, const double* ys, const double* ms, int64_t nnodes,
                       const int32_t* level, const uint32_t* first, const uint32_t* count,
                       const uint64_t* codes_sorted, int bits, double minx, double miny,
                       double root_side, const double* cx, const double* cy, const double* mass,
                       double* f) {
    for (int64_t k = 0; k < nnodes; ++k) {
        const double M = mass[k], l = ldexp(root_side, -level[k]);
        double sxx = 0, sxy = 0, syy = 0, t0 = 0, t1 = 0, t2 = 0, t3 = 0;
        for (uint32_t j = first[k]; j < first[k] + count[k]; ++j) {
            const double u = xs[j] - cx[k], v = ys[j] - cy[k], m = ms ? ms[j] : 1.0;
            sxx += m * u * u;
            sxy += m * u * v;
            syy += m * v * v;
            t0 += m * u * u * u;
            t1 += m * u * u * v;
            t2 += m * u * v * v;
            t3 += m * v * v * v;
        }
        if (count[k] > ORC_LEARN_T3_MAX) t0 = t1 = t2 = t3 = 0.0;
        const double a = 1.0 / (M * l * l), b = a / l;
        uint64_t ix, iy;
        deinterleave_prefix(codes_sorted[first[k]], bits, level[k], &ix, &iy);
        const double bx = minx + ((double)ix + 0.5) * l, by = miny + ((double)iy + 0.5) * l;
        double* o = f + k * ORC_CELL_FEATURES;
        o[0] = sxx * a; o[1] = sxy * a; o[2] = syy * a;
        o[3] = t0 * b; o[4] = t1 * b; o[5] = t2 * b; o[6] = t3 * b;
        o[7] = (cx[k] - bx) / l;
        o[8] = (cy[k] - by) / l;
        o[9] = log2((double)count[k]) / 16.0;
    }
}

void orc_learned_basis(double ux, double uy, double q, double* b) {
    const double w2r = ux * ux - uy * uy, w2i = 2.0 * ux * uy;
    const double w3r = w2r * ux - w2i * uy, w3i = w2r * uy + w2i * ux;
    const double w4r = w2r * w2r - w2i * w2i, w4i = 2.0 * w2r * w2i;
    b[0] = ux; b[1] = uy;            /* w */
    b[2] = ux; b[3] = -uy;           /* conj w */
    b[4] = w3r; b[5] = w3i;          /* w^3 */
    b[6] = q; b[7] = 0.0;            /* q */
    b[8] = q * w2r; b[9] = q * w2i;  /* q w^2 */
    b[10] = q * w2r; b[11] = -q * w2i; /* q conj(w)^2 */
    b[12] = q * w4r; b[13] = q * w4i;  /* q w^4 */
}

/* complex sum_j c_j b_j of 7 coefficient / basis pairs (re, im interleaved) */
static void learned_apply(const double* c, const double* b, double* re, double* im) {
    double x = 0, y = 0;
    for (int j = 0; j < ORC_LEARNED_TERMS; ++j) {
        x += c[2 * j] * b[2 * j] - c[2 * j + 1] * b[2 * j + 1];
        y += c[2 * j] * b[2 * j + 1] + c[2 * j + 1] * b[2 * j];
    }
    *re = x;
    *im = y;
}

void orc_pair_label(const double* xs, const double* ys, const double* ms, uint32_t first,
                    uint32_t count, double cx, double cy, double mass, double l, double px,
                    double py, int64_t self, double eps2, double* out) {
    double sx = 0, sy = 0;
    for (uint32_t j = first; j < first + count; ++j) {
        if ((int64_t)j == self) continue;
        const double dx = xs[j] - px, dy = ys[j] - py;
        const double inv = 1.0 / sqrt(dx * dx + dy * dy + eps2);
        const double w = (ms ? ms[j] : 1.0) * inv * inv * inv;
        sx += w * dx;
        sy += w * dy;
    }
    const double dx = cx - px, dy = cy - py, r2 = dx * dx + dy * dy;
    const double inv = 1.0 / sqrt(r2 + eps2);
    const double w = mass * inv * inv * inv;
    const double r = sqrt(r2), scale = mass * l * l / (r2 * r2);
    out[0] = (sx - w * dx) / scale;  /* label: residual in units of M l^2 / r^4 */
    out[1] = (sy - w * dy) / scale;
    out[2] = dx / r;                 /* offset features: direction and l / r */
    out[3] = dy / r;
    out[4] = l / r;
    out[5] = scale;
}

double orc_learned_loss_grad(const double* X, const double* T, int64_t B, int h, int terms,
                             const double* W1, const double* b1, const double* W2,
                             const double* b2, const double* W3, const double* b3, double* gW1,
                             double* gb1, double* gW2, double* gb2, double* gW3, double* gb3) {
    const int din = ORC_CELL_FEATURES, dout = 2 * ORC_LEARNED_TERMS, dx_ = din + 3;
    memset(gW1, 0, sizeof(double) * din * h);
    memset(gb1, 0, sizeof(double) * h);
    memset(gW2, 0, sizeof(double) * h * h);
    memset(gb2, 0, sizeof(double) * h);
    memset(gW3, 0, sizeof(double) * h * dout);
    memset(gb3, 0, sizeof(double) * dout);
    double* buf = (double*)malloc(sizeof(double) * (6 * h + 2 * dout));
    double *z1 = buf, *a1 = buf + h, *z2 = buf + 2 * h, *a2 = buf + 3 * h, *d2 = buf + 4 * h,
           *d1 = buf + 5 * h, *y = buf + 6 * h, *dy = buf + 6 * h + dout;
    double loss = 0.0, bas[2 * ORC_LEARNED_TERMS];
    for (int64_t s = 0; s < B; ++s) {
        const double* xin = X + s * dx_;
        mlp_fwd_one(xin, din, h, dout, W1, b1, W2, b2, W3, b3, z1, a1, z2, a2, y);
        orc_learned_basis(xin[din], xin[din + 1], xin[din + 2], bas);
        for (int j = terms; j < ORC_LEARNED_TERMS; ++j) bas[2 * j] = bas[2 * j + 1] = 0.0;
        double pr, pi;
        learned_apply(y, bas, &pr, &pi);
        const double ex = pr - T[2 * s], ey = pi - T[2 * s + 1];
        loss += (ex * ex + ey * ey) / (2.0 * (double)B);
        for (int j = 0; j < ORC_LEARNED_TERMS; ++j) {
            const double br = bas[2 * j], bi = bas[2 * j + 1];
            dy[2 * j] = (ex * br + ey * bi) / (double)B;
            dy[2 * j + 1] = (ey * br - ex * bi) / (double)B;
        }
        mlp_bwd_one(xin, din, h, dout, W2, W3, z1, a1, z2, a2, dy, d1, d2, gW1, gb1, gW2, gb2,
                    gW3, gb3);
    }
    free(buf);
    return loss;
}

/* the single-precision-decision walk of orc_bh_traverse_f32 (gravity) with the correction
 * s * sum_j c_j b_j added to every accepted cell with d2 <= l2 * inv_band2 in single precision
 * (l/r >= band; inv_band2 = 0: every accepted cell), c_j from coef[k * 14 ..] (FP64); the
 * correction is evaluated in FP64 from the single-precision node records and positions. */
void orc_bh_traverse_learned(const float* xs, const float* ys, const float* ms, int64_t n,
                             const int32_t* level, const uint32_t* first, const uint32_t* count,
                             const uint32_t* skip, const uint8_t* pp, const float* cx,
                             const float* cy, const float* mass, int64_t nnodes, float root_side,
                             float inv_theta2, double G, double eps2, const double* coef,
                             float inv_band2, const int64_t* sinks, int64_t nsinks,
                             double* out_ax, double* out_ay, uint64_t* counters) {
    (void)n;
    uint64_t pn = 0, ppc = 0, corrected = 0;
    double bas[2 * ORC_LEARNED_TERMS];
    for (int64_t s = 0; s < nsinks; ++s) {
        const int64_t i = sinks ? sinks[s] : s;
        const float px = xs[i], py = ys[i];
        double ax = 0, ay = 0, tot = 0;
        int64_t k = 0;
        while (k < nnodes) {
            const float sl = ldexpf(root_side, -level[k]);
            const float l2 = sl * sl;
            const float lim = l2 * inv_theta2;
            const float dx = cx[k] - px, dy = cy[k] - py;
            const float d2 = fmaf(dx, dx, dy * dy);
            if (d2 > lim) {
                interact(ORC_KERNEL_GRAVITY, 1.0, eps2, px, py, cx[k], cy[k], mass[k], &ax, &ay,
                         &tot);
                ++pn;
                const double ddx = (double)cx[k] - px, ddy = (double)cy[k] - py;
                const double r2 = ddx * ddx + ddy * ddy, r = sqrt(r2), l = (double)sl;
                if (inv_band2 == 0.0f || !(d2 > l2 * inv_band2)) {
                    orc_learned_basis(ddx / r, ddy / r, l / r, bas);
                    double cr, ci;
                    learned_apply(coef + k * 2 * ORC_LEARNED_TERMS, bas, &cr, &ci);
                    const double sc = (double)mass[k] * l * l / (r2 * r2);
                    ax += sc * cr;
                    ay += sc * ci;
                    ++corrected;
                }
                k = skip[k];
            } else if (pp[k]) {
                const int64_t f0 = first[k], e = f0 + count[k];
                for (int64_t q = f0; q < e; ++q) {
                    if (q == i) continue;
                    interact(ORC_KERNEL_GRAVITY, 1.0, eps2, px, py, xs[q], ys[q],
                             ms ? ms[q] : 1.0, &ax, &ay, &tot);
                    ++ppc;
                }
                k = skip[k];
            } else {
                ++k;
            }
        }
        out_ax[s] = G * ax;
        out_ay[s] = G * ay;
    }
    if (counters) {
        counters[0] = pn;
        counters[1] = ppc;
        counters[2] = corrected;
    }
}
