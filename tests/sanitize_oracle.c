/* Drives every entry point of the CPU oracle (oracle/bh_oracle.c) on small inputs, for builds
 * with AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_sanitizers.py). Prints "OK"
 * when every call returned; a sanitizer report aborts the run first. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../oracle/bh_oracle.h"

static uint64_t s_state = 0x243F6A8885A308D3ull;
static double uni(void) { /* splitmix64 -> [0, 1) */
    uint64_t z = (s_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

#define NEW(T, k) ((T*)calloc((size_t)(k) > 0 ? (size_t)(k) : 1, sizeof(T)))

static void run(int64_t n, int clustered) {
    double *x = NEW(double, n), *y = NEW(double, n), *m = NEW(double, n);
    float *xf = NEW(float, n), *yf = NEW(float, n), *mf = NEW(float, n);
    for (int64_t i = 0; i < n; ++i) {
        double a = uni(), b = uni();
        if (clustered && i % 3 == 0) a = b = 0.25; /* coincident points */
        x[i] = (double)(float)(2.0 * a - 1.0);
        y[i] = (double)(float)(2.0 * b - 1.0);
        m[i] = 1.0 / (double)n;
        xf[i] = (float)x[i];
        yf[i] = (float)y[i];
        mf[i] = (float)m[i];
    }
    double minx = x[0], miny = y[0], maxx = x[0], maxy = y[0];
    for (int64_t i = 1; i < n; ++i) {
        minx = fmin(minx, x[i]); miny = fmin(miny, y[i]);
        maxx = fmax(maxx, x[i]); maxy = fmax(maxy, y[i]);
    }
    const double L = fmax(maxx - minx, maxy - miny) > 0 ? fmax(maxx - minx, maxy - miny) : 1.0;

    uint32_t* c32 = NEW(uint32_t, n);
    uint64_t* c64 = NEW(uint64_t, n);
    uint64_t* q = NEW(uint64_t, n);
    uint32_t* perm = NEW(uint32_t, n);
    orc_morton32(x, y, n, minx, miny, L, c32);
    orc_morton64(x, y, n, minx, miny, L * 1.0001, 32, c64);
    orc_quantize_axis(x, n, minx, L, 20, q);
    orc_gpuref_codes32(xf, yf, n, 6, c32);
    orc_morton32(x, y, n, minx, miny, L, c32);
    orc_stable_sort_u64(c64, n, perm);

    double *xs = NEW(double, n), *ys = NEW(double, n), *ms = NEW(double, n);
    float *xsf = NEW(float, n), *ysf = NEW(float, n), *msf = NEW(float, n);
    uint64_t* cs = NEW(uint64_t, n);
    for (int64_t i = 0; i < n; ++i) {
        xs[i] = x[perm[i]]; ys[i] = y[perm[i]]; ms[i] = m[perm[i]];
        xsf[i] = (float)xs[i]; ysf[i] = (float)ys[i]; msf[i] = (float)ms[i];
        cs[i] = c64[perm[i]];
    }
    /* adaptive pre-order tree + moments + both walks */
    const int64_t cap = 2 * n + 64 * 33;
    int32_t* level = NEW(int32_t, cap);
    uint32_t *first = NEW(uint32_t, cap), *count = NEW(uint32_t, cap), *skip = NEW(uint32_t, cap),
             *nchild = NEW(uint32_t, cap);
    const int64_t nn = orc_adaptive_tree(cs, n, 32, 16, cap, level, first, count, skip, nchild);
    if (nn < 0) { printf("adaptive tree capacity\n"); exit(1); }
    uint8_t* pp = NEW(uint8_t, nn);
    for (int64_t k = 0; k < nn; ++k) pp[k] = nchild[k] <= 1;
    double *cx = NEW(double, nn), *cy = NEW(double, nn), *mass = NEW(double, nn),
           *m2 = NEW(double, 3 * nn);
    orc_node_moments(xs, ys, ms, nn, first, count, cx, cy, mass, m2);
    float *cxf = NEW(float, nn), *cyf = NEW(float, nn), *massf = NEW(float, nn);
    for (int64_t k = 0; k < nn; ++k) {
        cxf[k] = (float)cx[k]; cyf[k] = (float)cy[k]; massf[k] = (float)mass[k];
    }
    double *ax = NEW(double, n + 3), *ay = NEW(double, n + 3), *tot = NEW(double, n + 3);
    double *bx = NEW(double, n + 3), *by = NEW(double, n + 3);
    uint64_t cnt[3];
    int64_t sinks[3] = {0, n / 2, n - 1};
    for (int kernel = 0; kernel < 2; ++kernel) {
        orc_bh_traverse(xs, ys, ms, n, level, first, count, skip, pp, cx, cy, mass, nn, L * 1.0001,
                        ORC_SEM_QUADTREE, kernel, 1.0, 1e-6, 0.5, NULL, 0, ax, ay, tot, cnt, 3);
        orc_bh_traverse(xs, ys, ms, n, level, first, count, skip, pp, cx, cy, mass, nn, L * 1.0001,
                        ORC_SEM_MORTON, kernel, 1.0, 1e-6, 0.7, sinks, 3, ax, ay, tot, cnt, 1);
        orc_bh_traverse_f32(xsf, ysf, msf, n, level, first, count, skip, pp, cxf, cyf, massf, nn,
                            (float)(L * 1.0001), 4.0f, kernel, 1.0, 1e-6, cx, cy, L * 1.0001, 0.5,
                            ORC_SEM_QUADTREE, NULL, 0, ax, ay, tot, bx, by, cnt, 3);
        orc_direct(x, y, m, n, kernel, 1.0, 1e-6, sinks, 3, ax, ay, tot, 2);
    }
    /* fixed-depth Morton tree, its pre-order form, FMM */
    const int depth = orc_get_depth(16, n) < 1 ? 1 : orc_get_depth(16, n);
    const int64_t slots = orc_fixed_tree_slots(depth);
    uint32_t *fc = NEW(uint32_t, slots), *fp = NEW(uint32_t, slots), *fcn = NEW(uint32_t, slots);
    double *bbx = NEW(double, slots), *bby = NEW(double, slots), *bbl = NEW(double, slots),
           *cmx = NEW(double, slots), *cmy = NEW(double, slots), *m0 = NEW(double, slots),
           *fm2 = NEW(double, 3 * slots);
    orc_stable_sort_u64(cs, n, perm); /* already sorted: identity */
    uint32_t* c32s = NEW(uint32_t, n);
    {
        uint32_t* cc = NEW(uint32_t, n);
        uint64_t* cw = NEW(uint64_t, n);
        orc_morton32(x, y, n, minx, miny, L, cc);
        for (int64_t i = 0; i < n; ++i) cw[i] = cc[i];
        orc_stable_sort_u64(cw, n, perm);
        for (int64_t i = 0; i < n; ++i) {
            xs[i] = x[perm[i]]; ys[i] = y[perm[i]]; c32s[i] = cc[perm[i]];
        }
        free(cc);
        free(cw);
    }
    orc_fixed_tree(xs, ys, n, depth, minx, miny, maxx, maxy, fc, fp, fcn, bbx, bby, bbl, cmx, cmy,
                   m0, fm2);
    const int64_t cap2 = (int64_t)(depth + 1) * n + 1;
    int32_t* lv2 = NEW(int32_t, cap2);
    uint32_t *f2 = NEW(uint32_t, cap2), *n2 = NEW(uint32_t, cap2), *s2 = NEW(uint32_t, cap2),
             *h2 = NEW(uint32_t, cap2);
    orc_fixed_preorder(c32s, n, depth, cap2, lv2, f2, n2, s2, h2);
    double *fx = NEW(double, n), *fy = NEW(double, n);
    orc_fmm_sym(xs, ys, n, depth, fc, fp, fcn, bbl, cmx, cmy, m0, fm2, 0.5, fx, fy);

    /* QuadTree restatement */
    orc_quadtree* t = orc_quadtree_build(x, y, n, 16);
    const int64_t qn = orc_quadtree_node_count(t);
    int32_t* qd = NEW(int32_t, qn);
    uint32_t *qc = NEW(uint32_t, qn), *qh = NEW(uint32_t, qn), *qf = NEW(uint32_t, qn),
             *occ = NEW(uint32_t, n);
    double *lx = NEW(double, qn), *ly = NEW(double, qn), *hx = NEW(double, qn),
           *hy = NEW(double, qn), *qm = NEW(double, qn), *qx = NEW(double, qn),
           *qy = NEW(double, qn);
    orc_quadtree_export(t, qd, qc, qh, lx, ly, hx, hy, qm, qx, qy);
    orc_quadtree_occupants(t, qf, occ);
    memset(ax, 0, sizeof(double) * n);
    memset(ay, 0, sizeof(double) * n);
    orc_quadtree_forces(t, x, y, ORC_KERNEL_TSNE, 1.0, 1.0, 0.0, 0.5, ax, ay, cnt, 2);
    orc_quadtree_free(t);

    /* integrators and t-SNE */
    double *vx = NEW(double, n), *vy = NEW(double, n);
    double *px = NEW(double, n), *py = NEW(double, n);
    memcpy(px, x, sizeof(double) * n);
    memcpy(py, y, sizeof(double) * n);
    orc_leapfrog(px, py, vx, vy, m, n, 2, 1e-3, 1.0, 1e-6, 0.5, ORC_SEM_QUADTREE, 16, 32, 2);
    orc_leapfrog(px, py, vx, vy, m, n, 1, 1e-3, 1.0, 1e-6, 0.5, ORC_SEM_MORTON, 16, 16, 1);
    (void)orc_tsne_theta(0.2, 0.8, 17);
    uint32_t* rp = NEW(uint32_t, n + 1);
    uint32_t* col = NEW(uint32_t, 2 * n);
    double* val = NEW(double, 2 * n);
    int64_t nnz = 0;
    for (int64_t i = 0; i < n; ++i) {
        rp[i] = (uint32_t)nnz;
        if (n > 1) {
            col[nnz] = (uint32_t)((i + 1) % n);
            val[nnz++] = 0.5 / (double)n;
            col[nnz] = (uint32_t)((i + n - 1) % n);
            val[nnz++] = 0.5 / (double)n;
        }
    }
    rp[n] = (uint32_t)nnz;
    double *dx = NEW(double, n), *dy = NEW(double, n), *nx = NEW(double, n), *ny = NEW(double, n);
    orc_tsne_attraction(x, y, n, rp, col, val, 4.0, dx, dy);
    (void)orc_tsne_z(x, y, n, 2);
    (void)orc_tsne_kl(x, y, n, rp, col, val, 2);
    (void)orc_tsne_step(x, y, x, y, n, rp, col, val, 0.5, 16.0, 10.0, 0.5, dx, dy, nx, ny, 2);

    /* MLP and Adam */
    enum { DIN = 10, H = 8, DOUT = 2, B = 5 };
    double X[B * DIN], T[B * DOUT], Y[B * DOUT];
    double W1[DIN * H], b1[H], W2[H * H], b2[H], W3[H * DOUT], b3[DOUT];
    double g1[DIN * H], gb1[H], g2[H * H], gb2[H], g3[H * DOUT], gb3[DOUT];
    double* all[] = {X, T, W1, b1, W2, b2, W3, b3};
    const int sizes[] = {B * DIN, B * DOUT, DIN * H, H, H * H, H, H * DOUT, DOUT};
    for (int a = 0; a < 8; ++a)
        for (int i = 0; i < sizes[a]; ++i) all[a][i] = uni() - 0.5;
    orc_mlp_forward(X, B, DIN, H, DOUT, W1, b1, W2, b2, W3, b3, Y);
    (void)orc_mlp_loss_grad(X, T, B, DIN, H, DOUT, W1, b1, W2, b2, W3, b3, g1, gb1, g2, gb2, g3,
                            gb3);
    double mom1[DIN * H] = {0}, mom2[DIN * H] = {0};
    orc_adam(W1, g1, mom1, mom2, DIN * H, 1e-3, 0.9, 0.999, 1e-8, 1);

    void* blocks[] = {x, y, m, xf, yf, mf, c32, c64, q, perm, xs, ys, ms, xsf, ysf, msf, cs,
                      level, first, count, skip, nchild, pp, cx, cy, mass, m2, cxf, cyf, massf,
                      ax, ay, tot, bx, by, fc, fp, fcn, bbx, bby, bbl, cmx, cmy, m0, fm2, c32s,
                      lv2, f2, n2, s2, h2, fx, fy, qd, qc, qh, qf, occ, lx, ly, hx, hy, qm, qx, qy,
                      vx, vy, px, py, rp, col, val, dx, dy, nx, ny};
    for (size_t k = 0; k < sizeof(blocks) / sizeof(blocks[0]); ++k) free(blocks[k]);
}

int main(void) {
    const int64_t sizes[] = {1, 2, 17, 300, 2000};
    for (size_t k = 0; k < sizeof(sizes) / sizeof(sizes[0]); ++k) {
        run(sizes[k], 0);
        run(sizes[k], 1);
    }
    printf("OK\n");
    return 0;
}
