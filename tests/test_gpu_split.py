"""Split traversal (bh_set_traverse_split): F waves share a sink group and take its path
through the tree; wave f adds the monopoles of every F-th visit and the f-th slice of the sources
of every particle-particle cell, and the F partial results are summed in wave order.

Every (sink, node) interaction is evaluated by exactly one wave, so the interaction counters
must equal those of the one-wave walk exactly, and the forces must agree with it up to FP32
summation order. The oracle's forces are checked with the same tolerance as the one-wave walk
(median 1e-5, p99 1e-4 relative).
"""
import numpy as np
import pytest

from masterthesislucas_amd import BHContext, tsne
from masterthesislucas_amd._lib import BHError
from masterthesislucas_amd.workloads import plummer, two_plummer_disks, uniform_disk
from oracle import oracle as O
from tests.helpers import check_exact_walk, oracle_forces, oracle_tree, rel_err

pytestmark = pytest.mark.gpu

EPS = float(np.float32(1e-3))


def _forces(p, kernel, variant, split, theta, m=True, exact=False):
    ctx = BHContext(p.n, kernel=kernel, tree_mode="quadtree", softening=EPS)
    ctx.set_particles(p.x, p.y, p.m if m else None)
    ctx.set_variant(variant)
    ctx.set_split(split)
    ctx.build()
    ax, ay, tot = ctx.forces(theta)
    cnt = ctx.counters()
    if exact:
        check_exact_walk(ctx, p.x, p.y, p.m if m else None, kernel, theta,
                         EPS if kernel == "gravity" else 0.0, ax, ay, tot)
    ctx.close()
    return ax, ay, tot, cnt


@pytest.mark.parametrize("variant", [0, 4, 9])
@pytest.mark.parametrize("split", [2, 3, 8])
def test_split_equals_one_wave_walk(gpu, variant, split):
    p = two_plummer_disks(30000, seed=61)
    ax1, ay1, t1, c1 = _forces(p, "gravity", variant, 1, 0.5)
    ax, ay, t, c = _forces(p, "gravity", variant, split, 0.5)
    assert c == c1  # each interaction counted by exactly one wave
    r = rel_err(ax, ay, ax1.astype(np.float64), ay1.astype(np.float64))
    assert np.median(r) <= 1e-6 and r.max() <= 1e-4, (np.median(r), r.max())
    assert abs(t - t1) <= 1e-6 * abs(t1)


@pytest.mark.parametrize("kernel", ["gravity", "tsne"])
def test_split_matches_oracle(gpu, kernel):
    p = uniform_disk(20000, seed=62)
    grav = kernel == "gravity"
    ax, ay, tot, (pn, pp) = _forces(p, kernel, 4, 4, 0.5, m=grav, exact=True)
    t = oracle_tree(p.x, p.y, p.m if grav else None, mode="quadtree")
    okern = O.KERNEL_GRAVITY if grav else O.KERNEL_TSNE
    bx, by, bt, cnt = oracle_forces(t, okern, 0.5, eps=EPS if grav else 0.0)
    r = rel_err(ax, ay, bx, by)
    assert np.median(r) <= 1e-5 and np.quantile(r, 0.99) <= 1e-4
    assert abs(tot - bt.sum()) <= 1e-5 * abs(bt.sum())
    assert abs(int(pn) - int(cnt[0])) <= 1e-3 * cnt[0]
    assert abs(int(pp) - int(cnt[1])) <= 1e-3 * cnt[1]


@pytest.mark.parametrize("n", [1, 2, 17, 100, 1000])
def test_split_small_systems(gpu, n):
    """Cells with fewer sources than waves and walks with fewer visits than waves: some waves
    get empty slices and no monopoles and must write zero parts."""
    p = plummer(n, a=0.1, seed=63)
    ax1, ay1, _, c1 = _forces(p, "gravity", 4, 1, 0.5)
    ax, ay, _, c = _forces(p, "gravity", 4, 8, 0.5)
    assert c == c1
    assert np.allclose(ax, ax1, rtol=1e-5, atol=1e-6 * np.abs(ax1).max(initial=1.0))
    assert np.allclose(ay, ay1, rtol=1e-5, atol=1e-6 * np.abs(ay1).max(initial=1.0))


def test_split_setter_rejects_bad_values(gpu):
    ctx = BHContext(100, kernel="gravity")
    for bad in (0, 9, -1):
        with pytest.raises(BHError):
            ctx.set_split(bad)
    ctx.set_split(8)
    ctx.set_split(2)  # shrinking keeps the larger part buffer
    ctx.close()


def test_split_graph_steps_equal_eager_steps(gpu):
    p = plummer(20000, a=0.1, seed=64)
    ctxs = []
    for _ in range(2):
        c = BHContext(p.n, kernel="gravity", softening=EPS)
        c.set_particles(p.x, p.y, p.m, p.vx + 0.1 * p.y, p.vy - 0.1 * p.x)
        c.set_split(4)
        ctxs.append(c)
    eager, graph = ctxs
    for steps in (1, 2, 3):
        eager.leapfrog(steps, 1e-4, 0.5)
        graph.leapfrog_graph(steps, 1e-4, 0.5)
        for a, b in zip(eager.get_particles(), graph.get_particles()):
            assert np.array_equal(a, b), steps
    # changing the split re-captures the steps
    eager.set_split(2)
    graph.set_split(2)
    eager.leapfrog(2, 1e-4, 0.5)
    graph.leapfrog_graph(2, 1e-4, 0.5)
    for a, b in zip(eager.get_particles(), graph.get_particles()):
        assert np.array_equal(a, b)
    for c in ctxs:
        c.close()


def test_split_tsne_iterations_track_one_wave_walk(gpu):
    n = 4000
    X, _ = tsne.gaussian_mixture(n, dim=30, seed=65)
    rp, col, val = tsne.joint_probabilities(X, perplexity=20.0, k=60)
    yx, yy = tsne.initial_embedding(n, seed=65, std=1.0)
    runs = []
    for split in (1, 4):
        t = tsne.TSNE(rp, col, val, (yx, yy), min_theta=0.5, max_theta=0.5)
        t.ctx.set_split(split)
        try:
            dx, dy, z = t.timeStep(return_derivative=True)
            runs.append((dx, dy, z, t.embeddedPoints.copy()))
            # per-particle reads through the t-SNE context are sized to its n points
            ax, ay, _ = t.ctx.forces(0.5)
            assert ax.shape == (n,) and ay.shape == (n,)
        finally:
            t.close()
    (dx1, dy1, z1, y1), (dx4, dy4, z4, y4) = runs
    assert abs(z4 - z1) <= 1e-6 * z1
    scale = np.hypot(dx1, dy1).max()
    assert np.max(np.hypot(dx4 - dx1, dy4 - dy1)) <= 1e-5 * scale
    assert np.max(np.abs(y4 - y1)) <= 1e-5 * np.abs(y1).max()


@pytest.mark.parametrize("split", [2, 4, 8])
def test_split_crowded_leaf(gpu, split):
    """A leaf of 3,000 coincident points at the depth limit (as crowded as the leaves of a
    t-SNE embedding): its particle-particle loop is sliced over the waves; counts equal the
    one-wave walk's and forces agree with it up to summation order. (The per-cell sums of such
    a leaf are plain FP32 over 3,000 terms, beyond the 1e-5 bar of the compensated exact-walk
    check for the split and the one-wave walk alike.)"""
    rng = np.random.default_rng(64)
    x = rng.normal(0, 10, 5000).astype(np.float32)
    y = rng.normal(0, 10, 5000).astype(np.float32)
    x[:3000], y[:3000] = 1.25, -0.5
    from masterthesislucas_amd.workloads import Particles
    p = Particles(x, y, np.ones(5000, np.float32), np.zeros(5000, np.float32),
                  np.zeros(5000, np.float32))
    ax1, ay1, t1, c1 = _forces(p, "tsne", 21, 1, 0.5, m=False)
    ax, ay, t, c = _forces(p, "tsne", 21, split, 0.5, m=False)
    assert c == c1
    r = rel_err(ax, ay, ax1.astype(np.float64), ay1.astype(np.float64))
    assert np.median(r) <= 1e-6 and r.max() <= 1e-4, (np.median(r), r.max())
    assert abs(t - t1) <= 1e-6 * abs(t1)


@pytest.mark.parametrize("split", [2, 4])
def test_split_batched_systems(gpu, split):
    """Batched contexts: each work entry is walked by `split` waves. Counts equal the unsplit
    batched walk's and forces agree with it up to summation order, on untruncated Plummer
    systems whose one-child chains above the cluster make particle-particle cells of thousands
    of occupants."""
    from masterthesislucas_amd.workloads import batched_systems
    ps = batched_systems(8, seed=71)
    sizes = [len(q.x) for q in ps]
    cat = [np.concatenate([getattr(q, f) for q in ps]) for f in ("x", "y", "m")]
    res = []
    for F in (1, split):
        ctx = BHContext(int(sum(sizes)), kernel="gravity", softening=EPS, sizes=sizes)
        ctx.set_particles(*cat)
        ctx.set_variant(4)
        ctx.set_split(F)
        ctx.build()
        ax, ay, tot = ctx.forces(0.5)
        res.append((ax, ay, tot, ctx.counters()))
        ctx.close()
    (ax1, ay1, t1, c1), (ax, ay, t, c) = res
    assert c == c1 and c1[1] > 0
    # the particle-particle cells here hold thousands of sources summed in plain FP32 per cell:
    # a force that nearly cancels can move by a few 1e-4 of itself with the summation order
    r = rel_err(ax, ay, ax1.astype(np.float64), ay1.astype(np.float64))
    assert np.median(r) <= 1e-6 and np.quantile(r, 0.999) <= 1e-4 and r.max() <= 1e-3, (
        np.median(r), np.quantile(r, 0.999), r.max())
    assert abs(t - t1) <= 1e-6 * abs(t1)
