// C-ABI driver around the reference's own CPU Barnes-Hut — TEST INFRASTRUCTURE ONLY.
//
// Compiled by oracle/Makefile directly against the reference headers where they lie
// (nbodysolvers/cpu/nBodySolverBH.h, trees/cpu/quadtree.h, nbodysolvers/cpu/nBodySolverNaive.h,
// particles/tsnePoint2D.h); the output goes to oracle/_ref/. It is used to pin the C restatement
// (bh_oracle.c) and as the timed CPU baseline ("kind": "reference"). This file holds only the
// driver: the gravity kernels below are this project's (the reference has no live gravity
// kernel, nBodyInstances/gravitysim.h is commented out) and plug into the reference's
// std::function kernel slots (nBodySolverBH.h:20-31).
#include "particles/tsnePoint2D.h"
#include "nbodysolvers/cpu/nBodySolverBH.h"
#include "nbodysolvers/cpu/nBodySolverNaive.h"

#include <chrono>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

std::vector<TsnePoint2D> make_points(const double* x, const double* y, int64_t n) {
    std::vector<TsnePoint2D> pts;
    pts.reserve(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i)
        pts.emplace_back(glm::dvec2(x[i], y[i]), glm::dvec2(0.0), 0);
    return pts;
}

void walk(QuadTree<TsnePoint2D>* q, int depth, int64_t& k, int32_t* odepth, uint32_t* ocount,
          uint32_t* onchild, double* omass, double* ocomx, double* ocomy, double* olowx,
          double* olowy, double* ohighx, double* ohighy) {
    const int64_t me = k++;
    if (odepth) odepth[me] = depth;
    if (ocount) ocount[me] = static_cast<uint32_t>(q->occupants.size());
    if (onchild) onchild[me] = static_cast<uint32_t>(q->children.size());
    if (omass) omass[me] = static_cast<double>(q->totalMass);
    if (ocomx) ocomx[me] = q->centreOfMass.x;
    if (ocomy) ocomy[me] = q->centreOfMass.y;
    if (olowx) olowx[me] = q->lowestCorner.x;
    if (olowy) olowy[me] = q->lowestCorner.y;
    if (ohighx) ohighx[me] = q->highestCorner.x;
    if (ohighy) ohighy[me] = q->highestCorner.y;
    for (QuadTree<TsnePoint2D>* c : q->children)
        walk(c, depth + 1, k, odepth, ocount, onchild, omass, ocomx, ocomy, olowx, olowy, ohighx,
             ohighy);
}

int64_t count_nodes(QuadTree<TsnePoint2D>* q) {
    int64_t s = 1;
    for (QuadTree<TsnePoint2D>* c : q->children) s += count_nodes(c);
    return s;
}

}  // namespace

extern "C" {

// NBodySolverBH::updateTree + solveNbody on (x, y). kernel 0 = gravity (m per particle,
// G, eps2), 1 = the reference t-SNE kernels. Writes accelerations/derivatives (accumulated
// from zero), the build and solve wall times, and returns `total`.
double ref_bh_solve(const double* x, const double* y, int64_t n, int kernel, double G, double m,
                    double eps2, double theta, int max_children, double* out_ax, double* out_ay,
                    double* t_build, double* t_solve) {
    std::vector<TsnePoint2D> pts = make_points(x, y, n);
    std::function<void(double&, TsnePoint2D&, QuadTree<TsnePoint2D>*)> pn;
    std::function<void(double&, TsnePoint2D&, TsnePoint2D&)> pp;
    if (kernel == 1) {
        pn = &TSNEBHPNKernel;
        pp = &TSNEBHPPKernel;
    } else {
        pn = [G, m, eps2](double& total, TsnePoint2D& s, QuadTree<TsnePoint2D>* node) {
            const double M = static_cast<double>(node->totalMass) * m;
            const double dx = node->centreOfMass.x - s.position.x;
            const double dy = node->centreOfMass.y - s.position.y;
            const double r2 = dx * dx + dy * dy + eps2;
            const double inv = 1.0 / std::sqrt(r2);
            const double f = G * M * inv * inv * inv;
            s.derivative.x += f * dx;
            s.derivative.y += f * dy;
            total += -G * M * inv;
        };
        pp = [G, m, eps2](double& total, TsnePoint2D& s, TsnePoint2D& src) {
            const double dx = src.position.x - s.position.x;
            const double dy = src.position.y - s.position.y;
            const double r2 = dx * dx + dy * dy + eps2;
            const double inv = 1.0 / std::sqrt(r2);
            const double f = G * m * inv * inv * inv;
            s.derivative.x += f * dx;
            s.derivative.y += f * dy;
            total += -G * m * inv;
        };
    }
    NBodySolverBH<TsnePoint2D> solver(pn, pp, max_children, theta);
    auto t0 = Clock::now();
    solver.updateTree(pts, glm::dvec2(0.0), glm::dvec2(0.0));
    auto t1 = Clock::now();
    double total = 0.0;
    solver.solveNbody(total, pts);
    auto t2 = Clock::now();
    for (int64_t i = 0; i < n; ++i) {
        out_ax[i] = pts[i].derivative.x;
        out_ay[i] = pts[i].derivative.y;
    }
    if (t_build) *t_build = std::chrono::duration<double>(t1 - t0).count();
    if (t_solve) *t_solve = std::chrono::duration<double>(t2 - t1).count();
    return total;
}

// The reference QuadTree (pre-order, children in the reference's HH, HL, LH, LL order).
// Call with null arrays to get the node count.
int64_t ref_quadtree_export(const double* x, const double* y, int64_t n, int max_children,
                            int32_t* depth, uint32_t* count, uint32_t* nchild, double* mass,
                            double* comx, double* comy, double* lowx, double* lowy,
                            double* highx, double* highy) {
    std::vector<TsnePoint2D> pts = make_points(x, y, n);
    QuadTree<TsnePoint2D> root(max_children, &pts);
    if (!depth && !count && !nchild && !mass && !comx && !comy && !lowx && !lowy && !highx &&
        !highy)
        return count_nodes(&root);
    int64_t k = 0;
    walk(&root, 0, k, depth, count, nchild, mass, comx, comy, lowx, lowy, highx, highy);
    return k;
}

// NBodySolverNaive::solveNbody with TSNEnaiveKernel (exact t-SNE repulsion).
double ref_naive_tsne(const double* x, const double* y, int64_t n, double* out_ax, double* out_ay) {
    std::vector<TsnePoint2D> pts = make_points(x, y, n);
    NBodySolverNaive<TsnePoint2D> solver(&TSNEnaiveKernel);
    double total = 0.0;
    solver.solveNbody(total, pts);
    for (int64_t i = 0; i < n; ++i) {
        out_ax[i] = pts[i].derivative.x;
        out_ay[i] = pts[i].derivative.y;
    }
    return total;
}

}  // extern "C"
