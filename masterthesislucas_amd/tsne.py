"""t-SNE on the Barnes-Hut solver: input affinities and the iteration driver.

`TSNE` mirrors the reference driver `TSNE` (src/nBodyInstances/tsne.h:42-563): the same
learning rate N/15 (tsne.h:90), momentum 0.5/0.8 switching at iteration 250 (tsne.h:186),
exaggeration with the factor 4 of the gradient folded in (16 then 4, tsne.h:358), the cosine
theta schedule (tsne.h:208-233), the derivative (4/Z) * BH repulsion + sparse attraction
(tsne.h:272-301, 356-381), the momentum update (tsne.h:179-206) and the KL cost with log2
(tsne.h:383-460). All of an iteration runs on the GPU inside the library (include/bh.h,
bh_tsne_*); this class only holds parameters and the context.

The reference loads its P matrix from a Matrix Market file prepared offline with openTSNE
(MnistLoader.ipynb:539-547; tsne.h:94-98, Eigen loadMarket). Here the affinities come from
`affinities` / `joint_probabilities`, computed on the GPU by the library (bh_tsne_affinities):
exact k-nearest neighbours, a per-point Gaussian bandwidth found by bisection to the requested
perplexity, and the symmetrisation P = (P_cond + P_cond^T) / (2N), returned in CSR sorted by
(row, col) as setupEfficientPmatrix builds it (tsne.h:462-520).
"""
from __future__ import annotations

import ctypes as C
import numpy as np

from . import _lib
from .solver import BHContext, _ptr


# ---------------------------------------------------------------------------------------------
# input data and affinities
# ---------------------------------------------------------------------------------------------
def gaussian_mixture(n, dim=784, components=10, seed=4523, mean_std=2.0, sigma=1.0):
    """n points of an equal-weight isotropic Gaussian mixture: component means ~ N(0, mean_std^2
    I), points = mean + sigma N(0, I). Returns (X float32 [n, dim], labels uint8)."""
    rng = np.random.default_rng(seed)
    means = rng.normal(0.0, mean_std, size=(components, dim))
    labels = rng.integers(0, components, size=n).astype(np.uint8)
    X = means[labels] + rng.normal(0.0, sigma, size=(n, dim))
    return X.astype(np.float32), labels


def affinities(X, perplexity=30.0, k=90):
    """The t-SNE input affinities on the GPU (bh_tsne_affinities): exact k nearest neighbours,
    per-point bandwidths to the perplexity, and P = (P_cond + P_cond^T) / (2n).
    Returns dict(row_ptr uint32 [n+1], col uint32 [nnz], val float32 [nnz], idx uint32 [n, k],
    d2 float64 [n, k], beta float64 [n])."""
    X = np.ascontiguousarray(X, np.float32)
    n, dim = X.shape
    ctx = BHContext(max(int(n), 2), kernel="tsne")
    try:
        rp = np.empty(n + 1, np.uint32)
        col = np.empty(2 * n * k, np.uint32)
        val = np.empty(2 * n * k, np.float32)
        idx = np.empty((n, k), np.uint32)
        d2 = np.empty((n, k), np.float64)
        beta = np.empty(n, np.float64)
        nnz = C.c_uint32(0)
        ctx._c(_lib.load().bh_tsne_affinities(ctx.handle, _ptr(X), int(n), int(dim),
                                              float(perplexity), int(k), _ptr(rp), _ptr(col),
                                              _ptr(val), C.byref(nnz), _ptr(idx), _ptr(d2),
                                              _ptr(beta)))
    finally:
        ctx.close()
    m = nnz.value
    return dict(row_ptr=rp, col=col[:m].copy(), val=val[:m].copy(), idx=idx, d2=d2, beta=beta)


def joint_probabilities(X, perplexity=30.0, k=90, device=None):
    """Symmetrised sparse affinities of X computed on the GPU. Returns CSR (row_ptr uint32
    [n+1], col uint32 [nnz], val float32 [nnz]) sorted by (row, col), summing to 1. `device` is
    accepted for compatibility and ignored: the work runs on the library's device."""
    a = affinities(X, perplexity, k)
    return a["row_ptr"], a["col"], a["val"]


def initial_embedding(n, seed=4523, std=1e-2):
    """2-D start points ~ N(0, std^2 I) (the reference draws N(0, 0.01^2), tsne.h:530-548)."""
    rng = np.random.default_rng(seed + 1)
    Y = rng.normal(0.0, std, size=(n, 2))
    return Y[:, 0].astype(np.float32), Y[:, 1].astype(np.float32)


# ---------------------------------------------------------------------------------------------
# driver
# ---------------------------------------------------------------------------------------------
class TSNE:
    """Reference t-SNE driver over the GPU Barnes-Hut solver (tsne.h:42-563).

    `timeStep()` is one updateDerivative + updatePoints; `costFunction()` the KL divergence
    (FP64, exact Z); `embeddedPoints` the current positions (caller order, shape [N, 2]).
    """

    def __init__(self, row_ptr, col, val, y0, min_theta=0.5, max_theta=0.5,
                 exaggeration=(16.0, 4.0), momentum=(0.5, 0.8), switch=250, max_children=16,
                 learn_rate=None, solver="bh", _gpu_reference_depth=None):
        row_ptr = np.ascontiguousarray(row_ptr, np.uint32)
        col = np.ascontiguousarray(col, np.uint32)
        val = np.ascontiguousarray(val, np.float32)
        yx = np.ascontiguousarray(y0[0], np.float32)
        yy = np.ascontiguousarray(y0[1], np.float32)
        n = yx.size
        if row_ptr.size != n + 1:
            raise ValueError("row_ptr must hold n + 1 entries")
        self.n = n
        # "bh": Barnes-Hut over the adaptive QuadTree (NBodySolverBH); "fmm_sym": the symmetric
        # dual-tree FMM over the fixed-depth Morton tree (NBodySolverFMM_SYM_MORTON)
        mode = "morton_fixed" if solver == "fmm_sym" else "quadtree"
        depth = 0
        if _gpu_reference_depth is not None:
            mode, depth = "gpu_ref", int(_gpu_reference_depth)
        self.ctx = BHContext(n, kernel="tsne", max_children=max_children, theta=max_theta,
                             tree_mode=mode, fixed_depth=depth)
        if solver != "bh":
            self.ctx.set_solver(solver)
        L = _lib.load()
        p = _lib.TsneParams()
        if _gpu_reference_depth is not None:
            self.ctx._c(L.bh_tsne_params_gpu_reference(C.byref(p)))
        else:
            self.ctx._c(L.bh_tsne_params_default(C.byref(p), n, float(max_theta)))
            p.theta_min, p.theta_max = float(min_theta), float(max_theta)
            p.exag_early, p.exag_late = map(float, exaggeration)
            p.momentum_early, p.momentum_late = map(float, momentum)
            p.momentum_switch = p.exag_switch = int(switch)
        if learn_rate is not None:
            p.learning_rate = float(learn_rate)
        self.params = p
        self.iteration_counter = 0
        self.ctx._c(L.bh_tsne_set_p(self.ctx.handle, n, _ptr(row_ptr), _ptr(col), _ptr(val),
                                    int(col.size)))
        self.ctx._c(L.bh_tsne_set_embedding(self.ctx.handle, _ptr(yx), _ptr(yy)))
        # the context now holds the n embedding points: its per-particle reads size to n
        self.ctx.n = n

    @property
    def learnRate(self):
        return self.params.learning_rate

    def theta(self, iteration=None):
        it = self.iteration_counter if iteration is None else iteration
        return float(_lib.load().bh_tsne_theta(C.byref(self.params), int(it)))

    def timeStep(self, return_derivative=False):
        """One iteration. Returns (dx, dy, Z) when return_derivative, else Z."""
        L = _lib.load()
        z = C.c_double(0.0)
        dx = dy = None
        if return_derivative:
            dx, dy = np.empty(self.n, np.float32), np.empty(self.n, np.float32)
        self.ctx._c(L.bh_tsne_step(self.ctx.handle, C.byref(self.params), self.iteration_counter,
                                   _ptr(dx), _ptr(dy), C.byref(z)))
        self.iteration_counter += 1
        return (dx, dy, z.value) if return_derivative else z.value

    def run(self, iterations):
        """`iterations` steps without host synchronisation in between."""
        L = _lib.load()
        self.ctx._c(L.bh_tsne_run(self.ctx.handle, C.byref(self.params), self.iteration_counter,
                                  int(iterations)))
        self.iteration_counter += int(iterations)

    def costFunction(self):
        """(KL divergence, exact Z), both FP64."""
        kl, z = C.c_double(0.0), C.c_double(0.0)
        self.ctx._c(_lib.load().bh_tsne_kl(self.ctx.handle, C.byref(kl), C.byref(z)))
        return kl.value, z.value

    @property
    def embeddedPoints(self):
        yx, yy = np.empty(self.n, np.float32), np.empty(self.n, np.float32)
        self.ctx._c(_lib.load().bh_tsne_get_embedding(self.ctx.handle, _ptr(yx), _ptr(yy)))
        return np.stack([yx, yy], 1)

    def save_state(self, path, labels=None):
        """Write the current embedding and iteration counter (formats.save_state). The previous
        embedding is not part of the dump: a run resumed from it restarts the momentum term,
        like a fresh ``TSNE`` started from these positions."""
        from . import formats
        formats.save_state(path, self.iteration_counter, self.embeddedPoints, None, labels)

    def load_state(self, path):
        """Continue from a state dump: embedding and iteration counter (momentum restarts)."""
        from . import formats
        s = formats.load_state(path)
        if s["y"].shape[0] != self.n:
            raise ValueError(f"state holds {s['y'].shape[0]} points, this run {self.n}")
        yx = np.ascontiguousarray(s["y"][:, 0])
        yy = np.ascontiguousarray(s["y"][:, 1])
        self.ctx._c(_lib.load().bh_tsne_set_embedding(self.ctx.handle, _ptr(yx), _ptr(yy)))
        self.iteration_counter = s["iteration"]
        return s

    def close(self):
        self.ctx.close()


class TsneGpu(TSNE):
    """The reference's GPU t-SNE (TsneGpu, nBodyInstances/tsneGpu.h, with NBodySolverGpuBH): the
    complete fixed-depth tree with the centred root and centre splits (depth 6 by default,
    tsneGpu.h:145), theta 1, learning rate 1000, momentum 0.2, no exaggeration, derivative
    (1/Z) rep + attraction (nBodySolverGpuBH.cu:643-758). Same methods as `TSNE`; the
    per-point work runs in this library's kernels (include/bh.h: BH_TREE_GPU_REF,
    bh_tsne_params_gpu_reference)."""

    def __init__(self, row_ptr, col, val, y0, depth=6, learn_rate=None):
        super().__init__(row_ptr, col, val, y0, learn_rate=learn_rate,
                         _gpu_reference_depth=depth)


# ---------------------------------------------------------------------------------------------
# accuracy of the repulsion
# ---------------------------------------------------------------------------------------------
def nmae(d_naive, d_approx):
    """Mean over the points of |d_naive - d_approx| / |d_naive| (L2 norms per point), the error
    measure of the reference's harness (TsneTest::getNMAE, tsneTests.h:1366-1386). Arrays of
    shape [n, 2]; FP64 accumulation."""
    a = np.asarray(d_naive, np.float64).reshape(-1, 2)
    b = np.asarray(d_approx, np.float64).reshape(-1, 2)
    if a.shape != b.shape or a.shape[0] == 0:
        raise ValueError("two non-empty [n, 2] arrays of the same shape")
    return float(np.mean(np.hypot(a[:, 0] - b[:, 0], a[:, 1] - b[:, 1]) /
                         np.hypot(a[:, 0], a[:, 1])))


def repulsion_nmae(ctx, theta, sinks=None):
    """NMAE of the context's repulsion against the exact sum, as the harness measures it
    (tsneTests.h:143-165): each derivative is the repulsion scaled by 4/Z of its own solver
    (TSNE::updateRepulsive, tsne.h:287-301). The approximate side is the context's solver (the
    Barnes-Hut walk or the symmetric FMM) on its current tree; the exact side is the FP64 direct
    sum on the GPU over every point (Z_naive = the sum of all pairs), its derivative compared at
    `sinks` (all points by default). Returns (nmae, Z_approx, Z_naive)."""
    n = ctx.n
    ax, ay, z_approx = ctx.forces(theta)
    every = np.arange(n, dtype=np.uint32)
    dx, dy, tot = ctx.direct_fp64(every)
    z_naive = float(np.sum(tot))
    sel = every if sinks is None else np.asarray(sinks, np.int64)
    naive = np.stack([dx[sel], dy[sel]], 1) * (4.0 / z_naive)
    approx = np.stack([ax[sel], ay[sel]], 1).astype(np.float64) * (4.0 / z_approx)
    return nmae(naive, approx), z_approx, z_naive
