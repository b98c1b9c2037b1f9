// C++ adapter: the reference's solver plug-in interface over the C ABI of include/bh.h.
//
// NBodySolverHipBH<T> derives from the reference's NBodySolver<T>
// (src/nbodysolvers/cpu/nBodySolver.h:9-23) and keeps its three virtual methods and public
// knobs (maxChildren, theta, cell_size), so a driver that owns solvers through
// std::map<std::string, NBodySolver<T>*> (src/nBodyInstances/tsne.h:52, 104-122) can register
// it next to "BH", "BHR", ... and select it with nBodySelect. Include the reference's
// nbodysolvers/cpu/nBodySolver.h (and the point type) before this header.
//
// Behaviour kept from the reference:
//   updateTree(points, minPos, maxPos)  copies positions (and masses when T has a `mass`
//     member, else 1 like the QuadTree's particle counts, quadtree.h:163) and builds the tree.
//     Caller bounds are used in BH_TREE_MORTON_FIXED mode (nBodySolverFMM_SYM_MORTON.h:79-95)
//     and ignored in BH_TREE_QUADTREE mode (nBodySolverBH.h:43-46).
//   solveNbody(total, points)  sets total (FP64; t-SNE: Z) and ADDS the interaction sums to
//     points[i].derivative (nBodySolverBH.h:33-41); the caller zeroes derivatives first
//     (tsne.h:281-285). Results come back in the caller's order; caller memory is never
//     permuted.
//   getNodesBufferData(level)  line segments of the cell boxes at `level` (-1: all levels),
//     two vertices per edge in the order and colours of QuadTree::getNodesBufferData
//     (quadtree.h:174-208).
// Differences: kernels are the library's compiled gravity and t-SNE kernels chosen at
// construction (the reference injects std::function kernels, nBodySolverBH.h:20-31);
// errors throw std::runtime_error with the library's message (the reference has no error
// path). Positions and forces cross the boundary in FP32.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "bh.h"

template <typename T>
class NBodySolverHipBH : public NBodySolver<T> {
public:
    // solver BH_SOLVER_FMM_SYM stands in for NBodySolverFMM_SYM_MORTON (BH_KERNEL_TSNE with
    // BH_TREE_MORTON_FIXED); the default is the Barnes-Hut walk (NBodySolverBH)
    NBodySolverHipBH(bh_kernel kernel, int initMaxChildren, double initTheta, uint32_t n_max,
                     float G = 1.0f, float softening = 0.0f,
                     bh_tree_mode mode = BH_TREE_QUADTREE, int solver = BH_SOLVER_BH) {
        this->maxChildren = initMaxChildren;
        this->theta = initTheta;
        bh_config cfg;
        bh_config_default(&cfg);
        cfg.n_max = n_max;
        cfg.kernel = kernel;
        cfg.tree_mode = mode;
        cfg.max_children = initMaxChildren;
        cfg.G = G;
        cfg.softening = softening;
        cfg.theta = initTheta;
        code_bits_ = mode == BH_TREE_QUADTREE ? cfg.code_bits : 16;
        mode_ = mode;
        check(bh_create(&ctx_, &cfg), "bh_create");
        if (solver != BH_SOLVER_BH) check(bh_set_solver(ctx_, solver), "bh_set_solver");
    }

    NBodySolverHipBH(const NBodySolverHipBH&) = delete;
    NBodySolverHipBH& operator=(const NBodySolverHipBH&) = delete;
    ~NBodySolverHipBH() override { bh_destroy(ctx_); }

    void updateTree(std::vector<T>& points, glm::dvec2 minPos, glm::dvec2 maxPos) override {
        const size_t n = points.size();
        x_.resize(n);
        y_.resize(n);
        m_.resize(n);
        for (size_t i = 0; i < n; ++i) {
            x_[i] = static_cast<float>(points[i].position.x);
            y_[i] = static_cast<float>(points[i].position.y);
            if constexpr (requires(const T& p) { p.mass; })
                m_[i] = static_cast<float>(points[i].mass);
            else
                m_[i] = 1.0f;
        }
        check(bh_set_particles(ctx_, static_cast<uint32_t>(n), x_.data(), y_.data(), m_.data(),
                               nullptr, nullptr),
              "bh_set_particles");
        const double bounds[4] = {minPos.x, minPos.y, maxPos.x, maxPos.y};
        const bool caller_bounds = minPos.x <= maxPos.x && minPos.y <= maxPos.y;
        check(bh_build(ctx_, caller_bounds ? bounds : nullptr), "bh_build");
        n_ = n;
    }

    void solveNbody(double& total, std::vector<T>& points) override {
        if (points.size() != n_)
            throw std::runtime_error("solveNbody: points differ from the last updateTree");
        ax_.resize(n_);
        ay_.resize(n_);
        check(bh_forces(ctx_, this->theta, ax_.data(), ay_.data(), &total), "bh_forces");
        for (size_t i = 0; i < n_; ++i) {
            points[i].derivative.x += static_cast<double>(ax_[i]);
            points[i].derivative.y += static_cast<double>(ay_[i]);
        }
    }

    std::vector<VertexPos2Col3> getNodesBufferData(int level) override {
        std::vector<VertexPos2Col3> out;
        uint32_t nn = 0, max_level = 0, overfull = 0;
        check(bh_tree_info(ctx_, &nn, &max_level, &overfull), "bh_tree_info");
        std::vector<int32_t> lvl(nn);
        std::vector<uint32_t> first(nn);
        check(bh_get_tree(ctx_, lvl.data(), first.data(), nullptr, nullptr, nullptr, nullptr,
                          nullptr, nullptr),
              "bh_get_tree");
        std::vector<uint64_t> codes(n_);
        check(bh_get_morton(ctx_, codes.data(), nullptr), "bh_get_morton");
        double b[4], side = 0.0;
        check(bh_get_bounds(ctx_, b, &side), "bh_get_bounds");
        if (mode_ == BH_TREE_GPU_REF) {  // root centred on the bounding box (FP32 midpoint)
            b[0] = static_cast<double>((static_cast<float>(b[0]) + static_cast<float>(b[2])) * 0.5f) -
                   0.5 * side;
            b[1] = static_cast<double>((static_cast<float>(b[1]) + static_cast<float>(b[3])) * 0.5f) -
                   0.5 * side;
        }
        static const std::array<glm::vec3, 7> colors = {
            glm::vec3(1, 1, 1), glm::vec3(1, 0, 0), glm::vec3(0, 1, 0), glm::vec3(0, 0, 1),
            glm::vec3(1, 1, 0), glm::vec3(0, 1, 1), glm::vec3(1, 0, 1)};
        const glm::vec3 color = colors[std::min(level + 1, 6)];
        for (uint32_t k = 0; k < nn; ++k) {
            const int l = lvl[k];
            if (level != -1 && l != level) continue;
            // cell of the node's first particle at level l: the top 2l bits of its code
            const uint64_t prefix = l ? codes[first[k]] >> (2 * (code_bits_ - l)) : 0;
            uint64_t ix = 0, iy = 0;
            for (int bit = 0; bit < l; ++bit) {
                ix |= ((prefix >> (2 * bit)) & 1u) << bit;
                iy |= ((prefix >> (2 * bit + 1)) & 1u) << bit;
            }
            const double cell = side / static_cast<double>(1ull << l);
            const float lx = static_cast<float>(b[0] + cell * static_cast<double>(ix));
            const float ly = static_cast<float>(b[1] + cell * static_cast<double>(iy));
            const float hx = static_cast<float>(b[0] + cell * static_cast<double>(ix + 1));
            const float hy = static_cast<float>(b[1] + cell * static_cast<double>(iy + 1));
            const glm::vec2 corners[8] = {{lx, ly}, {hx, ly}, {lx, ly}, {lx, hy},
                                          {lx, hy}, {hx, hy}, {hx, ly}, {hx, hy}};
            for (const glm::vec2& c : corners) out.push_back(VertexPos2Col3(c, color));
        }
        return out;
    }

    bh_ctx* context() { return ctx_; }

private:
    void check(bh_status st, const char* what) {
        if (st != BH_OK)
            throw std::runtime_error(std::string(what) + ": " + bh_last_error(ctx_));
    }

    bh_ctx* ctx_ = nullptr;
    int code_bits_ = 32;
    bh_tree_mode mode_ = BH_TREE_QUADTREE;
    size_t n_ = 0;
    std::vector<float> x_, y_, m_, ax_, ay_;
};
