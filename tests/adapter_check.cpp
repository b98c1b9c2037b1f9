// Drop-in check of include/nbody_solver_hip.hpp — TEST INFRASTRUCTURE ONLY.
//
// Built by `make -C oracle adapter` against the reference's own headers where they lie (the
// NBodySolver<T> base, TsnePoint2D, NBodySolverBH with its t-SNE kernels) into
// oracle/_ref/adapter_check, and run on a GPU by tests/test_gpu_adapter.py. The same
// std::vector<TsnePoint2D> goes through the reference's CPU solver and through
// NBodySolverHipBH, both held as NBodySolver<TsnePoint2D>* and called in the order of
// TSNE::updateRepulsive (src/nBodyInstances/tsne.h:250-301): zero derivatives, updateTree
// with the caller's bounds, solveNbody, scale by 4/Z. Prints one line of figures and exits
// non-zero when a tolerance is missed.
#include "particles/tsnePoint2D.h"
#include "nbodysolvers/cpu/nBodySolverBH.h"

#include "nbody_solver_hip.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

namespace {

// deterministic 2-D Gaussian blobs (splitmix64 + Box-Muller)
uint64_t g_state = 0x9E3779B97F4A7C15ull;
double uniform01() {
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (static_cast<double>(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
double normal() {
    return std::sqrt(-2.0 * std::log(uniform01())) * std::cos(6.283185307179586 * uniform01());
}

std::vector<TsnePoint2D> make_points(int n) {
    g_state = 0x9E3779B97F4A7C15ull;  // same points on every call
    std::vector<TsnePoint2D> pts;
    pts.reserve(n);
    for (int i = 0; i < n; ++i) {
        const int c = i % 5;
        // positions rounded to FP32 so both solvers see identical inputs
        const float x = static_cast<float>(3.0 * std::cos(1.3 * c) + 0.6 * normal());
        const float y = static_cast<float>(3.0 * std::sin(1.3 * c) + 0.4 * normal());
        pts.emplace_back(glm::dvec2(x, y), glm::dvec2(0.0), c);
    }
    return pts;
}

double quantile(std::vector<double> v, double q) {
    std::sort(v.begin(), v.end());
    return v[static_cast<size_t>(q * (v.size() - 1))];
}

}  // namespace

// `adapter_check N THETA fmm OUT`: the adapter with BH_SOLVER_FMM_SYM on a fixed-depth Morton
// tree, in the place tsne.h:121-122 registers NBodySolverFMM_SYM_MORTON (that solver needs
// boost's sort, absent here, so the comparison is made by the caller against the oracle).
// Writes n (int32), the FP32 positions x then y, the FP64 derivatives x then y and Z to OUT.
int fmm_dump(int n, double theta, const char* out) {
    NBodySolver<TsnePoint2D>* solver = new NBodySolverHipBH<TsnePoint2D>(
        BH_KERNEL_TSNE, 16, theta, n, 1.0f, 0.0f, BH_TREE_MORTON_FIXED, BH_SOLVER_FMM_SYM);
    std::vector<TsnePoint2D> pts = make_points(n);
    glm::dvec2 lo(1e300), hi(-1e300);
    for (auto& p : pts) {
        lo = glm::min(lo, p.position);
        hi = glm::max(hi, p.position);
        p.derivative = glm::dvec2(0.0);
    }
    solver->updateTree(pts, lo, hi);
    double total = 0.0;
    solver->solveNbody(total, pts);
    delete solver;
    std::FILE* f = std::fopen(out, "wb");
    if (!f) return 2;
    std::vector<float> pos(2 * static_cast<size_t>(n));
    std::vector<double> der(2 * static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        pos[i] = static_cast<float>(pts[i].position.x);
        pos[n + i] = static_cast<float>(pts[i].position.y);
        der[i] = pts[i].derivative.x;
        der[n + i] = pts[i].derivative.y;
    }
    const int32_t n32 = n;
    const bool ok = std::fwrite(&n32, sizeof n32, 1, f) == 1 &&
                    std::fwrite(pos.data(), sizeof(float), pos.size(), f) == pos.size() &&
                    std::fwrite(der.data(), sizeof(double), der.size(), f) == der.size() &&
                    std::fwrite(&total, sizeof total, 1, f) == 1;
    std::fclose(f);
    std::printf("{\"n\": %d, \"theta\": %g, \"z\": %.17g}\n", n, theta, total);
    return ok ? 0 : 2;
}

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 20000;
    const double theta = argc > 2 ? std::atof(argv[2]) : 0.5;
    if (argc > 4 && std::string(argv[3]) == "fmm") return fmm_dump(n, theta, argv[4]);
    std::map<std::string, NBodySolver<TsnePoint2D>*> solvers;
    solvers["BH"] = new NBodySolverBH<TsnePoint2D>(&TSNEBHPNKernel, &TSNEBHPPKernel, 16, theta);
    solvers["HIP_BH"] = new NBodySolverHipBH<TsnePoint2D>(BH_KERNEL_TSNE, 16, theta, n);

    std::map<std::string, std::vector<TsnePoint2D>> result;
    std::map<std::string, double> z;
    for (auto& [name, solver] : solvers) {
        std::vector<TsnePoint2D> pts = make_points(n);
        glm::dvec2 lo(1e300), hi(-1e300);
        for (auto& p : pts) {
            lo = glm::min(lo, p.position);
            hi = glm::max(hi, p.position);
            p.derivative = glm::dvec2(0.0);
        }
        solver->updateTree(pts, lo, hi);
        double total = 0.0;
        solver->solveNbody(total, pts);
        for (auto& p : pts) p.derivative *= 4.0 / total;
        result[name] = std::move(pts);
        z[name] = total;
    }

    std::vector<double> rel(n);
    for (int i = 0; i < n; ++i) {
        const glm::dvec2 a = result["HIP_BH"][i].derivative, b = result["BH"][i].derivative;
        rel[i] = glm::length(a - b) / std::max(glm::length(b), 1e-300);
    }
    const double med = quantile(rel, 0.5), p99 = quantile(rel, 0.99);
    const double zrel = std::fabs(z["HIP_BH"] - z["BH"]) / z["BH"];

    // cell boxes of level 3 as drawn by both solvers (vertex multisets)
    auto boxes = [](std::vector<VertexPos2Col3> v) {
        std::vector<std::tuple<float, float, float>> out;
        for (auto& e : v) out.emplace_back(e.position.x, e.position.y, e.color.r + 2 * e.color.g);
        std::sort(out.begin(), out.end());
        return out;
    };
    const auto rb = boxes(solvers["BH"]->getNodesBufferData(3));
    const auto hb = boxes(solvers["HIP_BH"]->getNodesBufferData(3));
    double box_err = rb.size() == hb.size() ? 0.0 : 1e30;
    for (size_t i = 0; i < rb.size() && i < hb.size(); ++i)
        box_err = std::max({box_err, (double)std::fabs(std::get<0>(rb[i]) - std::get<0>(hb[i])),
                            (double)std::fabs(std::get<1>(rb[i]) - std::get<1>(hb[i])),
                            (double)std::fabs(std::get<2>(rb[i]) - std::get<2>(hb[i]))});

    std::printf("{\"n\": %d, \"theta\": %g, \"median_rel\": %.3e, \"p99_rel\": %.3e, "
                "\"z_rel\": %.3e, \"boxes\": %zu, \"box_vertices_ref\": %zu, \"box_err\": %.3e}\n",
                n, theta, med, p99, zrel, hb.size(), rb.size(), box_err);
    for (auto& [name, solver] : solvers) delete solver;
    const bool ok = med <= 1e-5 && p99 <= 1e-4 && zrel <= 1e-6 && box_err <= 1e-5;
    return ok ? 0 : 1;
}
