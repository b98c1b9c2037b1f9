"""t-SNE iteration on the GPU (include/bh.h, bh_tsne_*) against the oracle's restatement of the
reference driver (tsne.h:179-233, 272-301, 356-460).

Tolerances: the GPU computes in FP32 with the Barnes-Hut repulsion on the same tree topology as
the reference QuadTree, the oracle in FP64. Per point, errors are measured against the scale
|4/Z rep_i| + |att_i| of the two terms that make up the derivative (their sum can cancel), with
the FP32 bar of the force tests: median 1e-5, p99 1e-4.
"""
import ctypes as C

import numpy as np
import pytest

from masterthesislucas_amd import _lib, tsne
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

MEDIAN_TOL, P99_TOL = 1e-5, 1e-4


def _problem(n, dim=40, seed=11, perplexity=20.0, k=60):
    X, _ = tsne.gaussian_mixture(n, dim=dim, seed=seed)
    rp, col, val = tsne.joint_probabilities(X, perplexity=perplexity, k=k)
    yx, yy = tsne.initial_embedding(n, seed=seed, std=1.0)
    return rp, col, val, yx, yy


def _oracle_terms(x, y, rp, col, val, theta, exag):
    qt = orc.QuadTree(x, y, 16)
    rx, ry, z = qt.forces(orc.KERNEL_TSNE, 1.0, 1.0, 0.0, theta, nthreads=8)[:3]
    ax, ay = orc.tsne_attraction(x, y, rp, col, val, exag)
    return 4.0 / z * rx, 4.0 / z * ry, ax, ay, z


def _scaled_err(got, want, scale):
    e = np.abs(got - want) / scale
    return float(np.median(e)), float(np.quantile(e, 0.99))


@pytest.mark.parametrize("n,theta", [(3000, 0.5), (6000, 0.8)])
def test_tsne_step_matches_oracle(gpu, n, theta):
    rp, col, val, yx, yy = _problem(n)
    t = tsne.TSNE(rp, col, val, (yx, yy), min_theta=theta, max_theta=theta)
    try:
        x0, y0 = yx.astype(np.float64), yy.astype(np.float64)
        # iteration 0: previous positions equal the current ones (tsne.h:553-555)
        dx, dy, z = t.timeStep(return_derivative=True)
        odx, ody, onx, ony, oz = orc.tsne_step(x0, y0, x0, y0, rp, col, val, theta, 16.0,
                                               n / 15.0, 0.5)
        rx, ry, ax, ay, _ = _oracle_terms(x0, y0, rp, col, val, theta, 16.0)
        scale = np.hypot(np.abs(rx) + np.abs(ax), np.abs(ry) + np.abs(ay))
        med, p99 = _scaled_err(np.hypot(dx - odx, dy - ody), 0.0, scale)
        assert med < MEDIAN_TOL and p99 < P99_TOL, (med, p99)
        assert abs(z - oz) < 1e-6 * oz
        y1 = t.embeddedPoints.astype(np.float64)
        step = (n / 15.0) * scale + 1e-30
        med, p99 = _scaled_err(np.hypot(y1[:, 0] - onx, y1[:, 1] - ony), 0.0,
                               step + np.hypot(x0, y0) * 1e-6)
        assert med < MEDIAN_TOL and p99 < P99_TOL, (med, p99)

        # iteration 1 from the GPU's own state: exercises the momentum term mu (y - y_prev)
        dx, dy, z = t.timeStep(return_derivative=True)
        x1, y1_ = y1[:, 0], y1[:, 1]
        odx, ody, onx, ony, oz = orc.tsne_step(x1, y1_, x0, y0, rp, col, val, theta, 16.0,
                                               n / 15.0, 0.5)
        rx, ry, ax, ay, _ = _oracle_terms(x1, y1_, rp, col, val, theta, 16.0)
        scale = np.hypot(np.abs(rx) + np.abs(ax), np.abs(ry) + np.abs(ay))
        med, p99 = _scaled_err(np.hypot(dx - odx, dy - ody), 0.0, scale)
        assert med < MEDIAN_TOL and p99 < P99_TOL, (med, p99)
        y2 = t.embeddedPoints.astype(np.float64)
        mom = 0.5 * np.hypot(x1 - x0, y1_ - y0)
        med, p99 = _scaled_err(np.hypot(y2[:, 0] - onx, y2[:, 1] - ony), 0.0,
                               (n / 15.0) * scale + mom + np.hypot(x1, y1_) * 1e-6)
        assert med < MEDIAN_TOL and p99 < P99_TOL, (med, p99)
    finally:
        t.close()


def test_tsne_late_phase_constants(gpu):
    # iteration >= 250: exaggeration 4, momentum 0.8 (tsne.h:186, 358)
    n, theta = 3000, 0.5
    rp, col, val, yx, yy = _problem(n, seed=12)
    t = tsne.TSNE(rp, col, val, (yx, yy), min_theta=theta, max_theta=theta)
    try:
        t.iteration_counter = 300
        dx, dy, _ = t.timeStep(return_derivative=True)
        x0, y0 = yx.astype(np.float64), yy.astype(np.float64)
        odx, ody, onx, ony, _ = orc.tsne_step(x0, y0, x0, y0, rp, col, val, theta, 4.0,
                                              n / 15.0, 0.8)
        rx, ry, ax, ay, _ = _oracle_terms(x0, y0, rp, col, val, theta, 4.0)
        scale = np.hypot(np.abs(rx) + np.abs(ax), np.abs(ry) + np.abs(ay))
        med, p99 = _scaled_err(np.hypot(dx - odx, dy - ody), 0.0, scale)
        assert med < MEDIAN_TOL and p99 < P99_TOL, (med, p99)
    finally:
        t.close()


def test_tsne_kl_matches_oracle(gpu):
    n = 4000
    rp, col, val, yx, yy = _problem(n, seed=13)
    t = tsne.TSNE(rp, col, val, (yx, yy))
    try:
        kl, z = t.costFunction()
        x, y = yx.astype(np.float64), yy.astype(np.float64)
        oz = orc.tsne_z(x, y)
        okl = orc.tsne_kl(x, y, rp, col, val.astype(np.float64))
        assert abs(z - oz) <= 1e-11 * oz
        assert abs(kl - okl) <= 1e-9 * abs(okl)
        assert kl > 0
    finally:
        t.close()


def test_tsne_run_descends_and_tracks_oracle(gpu):
    # 40 iterations, early exaggeration: the KL divergence falls; the GPU trajectory and the
    # oracle's stay close (a sanity bound only: FP32 and FP64 trajectories drift apart)
    n, theta = 2000, 0.5
    rp, col, val, yx, yy = _problem(n, seed=14)
    t = tsne.TSNE(rp, col, val, (yx, yy), min_theta=theta, max_theta=theta)
    try:
        kl0, _ = t.costFunction()
        t.run(40)
        kl1, _ = t.costFunction()
        assert kl1 < kl0
        x, y = yx.astype(np.float64), yy.astype(np.float64)
        px, py = x.copy(), y.copy()
        for it in range(40):
            _, _, nx, ny, _ = orc.tsne_step(x, y, px, py, rp, col, val, theta, 16.0, n / 15.0,
                                            0.5)
            px, py, x, y = x, y, nx, ny
        okl = orc.tsne_kl(x, y, rp, col, val.astype(np.float64))
        assert abs(kl1 - okl) < 1e-2 * okl, (kl1, okl)
    finally:
        t.close()


def test_tsne_rejects_bad_input(gpu):
    n = 100
    rp, col, val, yx, yy = _problem(n, seed=15, k=10, perplexity=5.0)
    bad = col.copy()
    bad[3] = n
    with pytest.raises(_lib.BHError):
        tsne.TSNE(rp, bad, val, (yx, yy))
    rp2 = rp.copy()
    rp2[-1] += 1
    with pytest.raises(_lib.BHError):
        tsne.TSNE(rp2, col, val, (yx, yy))


def test_direct_sum_between_steps_keeps_p_matrix(gpu):
    # regression: the FP64 direct sum grows its own buffers and must leave P untouched
    n = 2000
    rp, col, val, yx, yy = _problem(n, seed=16)
    a = tsne.TSNE(rp, col, val, (yx, yy))
    b = tsne.TSNE(rp, col, val, (yx, yy))
    try:
        a.timeStep()
        b.timeStep()
        a.ctx.direct_fp64(np.arange(0, n, 3, dtype=np.uint32))
        da = a.timeStep(return_derivative=True)
        db = b.timeStep(return_derivative=True)
        assert np.array_equal(da[0], db[0]) and np.array_equal(da[1], db[1]) and da[2] == db[2]
    finally:
        a.close()
        b.close()


def test_state_dump_resumes_the_run(gpu, tmp_path):
    """A run saved with TSNE.save_state and resumed in a new object continues from the same
    embedding and iteration (the momentum term restarts, as documented)."""
    rp, col, val, yx, yy = _problem(2000)
    a = tsne.TSNE(rp, col, val, (yx, yy), min_theta=0.5, max_theta=0.5)
    b = tsne.TSNE(rp, col, val, (yx, yy), min_theta=0.5, max_theta=0.5)
    try:
        a.run(30)
        a.save_state(tmp_path / "s.bin")
        s = b.load_state(tmp_path / "s.bin")
        assert b.iteration_counter == 30 == s["iteration"]
        assert np.array_equal(b.embeddedPoints, a.embeddedPoints)
        # from equal positions with the momentum reset on both, one step agrees
        a.load_state(tmp_path / "s.bin")
        da = a.timeStep(return_derivative=True)
        db = b.timeStep(return_derivative=True)
        assert np.array_equal(da[0], db[0]) and np.array_equal(a.embeddedPoints, b.embeddedPoints)
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("solver", ["bh", "fmm_sym"])
def test_repulsion_nmae_matches_oracle(gpu, solver):
    """The harness' error measure (tsneTests.h:143-165, getNMAE :1366-1386) on the GPU: both
    derivatives scaled by 4/Z of their own solver, compared over every point. Against the same
    measure from the oracle's solver (same tree and decisions) and its FP64 direct sum."""
    from masterthesislucas_amd import BHContext
    from tests.helpers import oracle_tree

    rng = np.random.default_rng(81)
    k = rng.integers(0, 5, 6000)
    p = rng.normal(size=(6000, 2)) * 0.4 + rng.normal(size=(5, 2))[k] * 6.0
    x, y = p[:, 0].astype(np.float32), p[:, 1].astype(np.float32)
    mode = "quadtree" if solver == "bh" else "morton_fixed"
    ctx = BHContext(x.size, kernel="tsne", tree_mode=mode)
    if solver != "bh":
        ctx.set_solver(solver)
    ctx.set_particles(x, y)
    ctx.build()
    e_gpu, z_gpu, z_naive = tsne.repulsion_nmae(ctx, 0.5)
    e_zero = tsne.repulsion_nmae(ctx, 1e-6)[0]
    ctx.close()
    ex, ey, et = orc.direct(x.astype(np.float64), y.astype(np.float64), None, orc.KERNEL_TSNE)
    assert z_naive == pytest.approx(et.sum(), rel=1e-11)
    t = oracle_tree(x, y, None, mode=mode)
    if solver == "bh":
        qt = orc.QuadTree(x.astype(np.float64), y.astype(np.float64), 16)
        bx, by, bz = qt.forces(orc.KERNEL_TSNE, 1.0, 1.0, 0.0, 0.5, nthreads=8)[:3]
    else:
        minx, miny, maxx, maxy = t["bounds"]
        tree = orc.fixed_tree(t["xs"], t["ys"], t["depth"], minx, miny, maxx, maxy)
        sx, sy, bz = orc.fmm_sym(t["xs"], t["ys"], tree, t["depth"], 0.5)
        bx, by = np.empty(x.size), np.empty(x.size)
        bx[t["perm"]] = sx
        by[t["perm"]] = sy
    e_orc = tsne.nmae(np.stack([ex, ey], 1) * (4.0 / et.sum()), np.stack([bx, by], 1) * (4.0 / bz))
    assert abs(e_gpu - e_orc) <= 1e-3 * e_orc + 1e-6, (e_gpu, e_orc)
    assert z_gpu == pytest.approx(bz, rel=1e-5)
    assert e_zero <= 1e-5  # theta -> 0: every pair direct, the exact sum up to FP32 output
