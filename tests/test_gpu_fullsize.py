"""The benchmark configurations at their full sizes, on sampled sinks.

Config 2: two Plummer disks, N = 1,048,576, theta in {0.3, 0.5, 0.7, 1.0}: 4,096 seeded sinks
  against the oracle's single-precision-decision walk over the GPU's node records (every sink
  within WALK_TOL of its |contribution| sums), the FP64-rule disagreements counted against the
  FP64 oracle on its own tree, and the Barnes-Hut error against the FP64 direct sum growing with
  theta.
Config 3: t-SNE, 70,000 points of the 10-component mixture in 784-D, perplexity 30, k = 90: one
  iteration against the oracle's restatement of the reference driver.
Config 4: 256 systems, N log-uniform in [1,024, 262,144] (11.1 M particles): 16 seeded sinks of
  every system against the oracle walk over that system's own subtree.
Config 5: Plummer a = 1, N = 4,194,304: a network trained on 2M pairs of a second realisation;
  on 4,096 seeded sinks, the learned walk at theta = 0.9 against the oracle's learned walk over
  the GPU's node records with the GPU's own coefficient records (exact bar), and its error
  against the FP64 direct sum below plain theta = 0.9's (a third of it) and plain theta = 0.5's.
"""
import numpy as np
import pytest

from masterthesislucas_amd import BHContext, tsne
from masterthesislucas_amd.workloads import (SEED_PLUMMER, batched_systems, plummer,
                                             sample_indices, two_plummer_disks)
from oracle import oracle as O
from tests.helpers import (WALK_TOL, batched_export, exact_walk, exact_walk_system, oracle_tree,
                           rel_err, walk_error)

pytestmark = pytest.mark.gpu

EPS = float(np.float32(1e-3))


def test_config2_two_disks_1m_sampled_sinks(gpu):
    n = 1 << 20
    p = two_plummer_disks(n)
    ctx = BHContext(n, kernel="gravity", tree_mode="quadtree", softening=EPS)
    ctx.set_particles(p.x, p.y, p.m)
    ctx.build()
    s = sample_indices(n, 4096)
    t = oracle_tree(p.x, p.y, p.m, mode="quadtree")
    where = np.empty(n, np.int64)
    where[t["perm"]] = np.arange(n)
    dx, dy, _ = ctx.direct_fp64(s)  # FP64 direct sum on the GPU (pinned against the CPU oracle)
    errs = []
    for theta in (0.3, 0.5, 0.7, 1.0):
        ax, ay, _ = ctx.forces(theta)
        w = exact_walk(ctx, p.x, p.y, p.m, "gravity", theta, eps=EPS, sinks=s, t64=t)
        e = walk_error(ax[s], ay[s], w)
        assert e.max() <= WALK_TOL, (theta, float(e.max()))
        bx, by, _, cnt = O.bh_traverse(t["xs"], t["ys"], t["ms"], t["tree"], t["mom"], t["side"],
                                       O.SEM_QUADTREE, O.KERNEL_GRAVITY, G=1.0, eps2=EPS * EPS,
                                       theta=theta, sinks=where[s], pp=t["pp"])
        r = rel_err(ax[s], ay[s], bx, by)
        err = rel_err(ax[s], ay[s], dx, dy)
        errs.append(float(np.median(err)))
        print(f"config 2 theta={theta}: FP64-rule disagreements on 4,096 walks {w['mismatch']} "
              f"of {w['pn'] + w['pp']}; vs FP64 oracle median {np.median(r):.2e}; "
              f"BH error vs direct sum median {errs[-1]:.3e}")
        assert w["mismatch"] <= 1e-4 * (w["pn"] + w["pp"])
        assert np.median(r) <= 1e-5
    assert errs == sorted(errs)
    ctx.close()


def test_config3_tsne_70k_one_iteration(gpu):
    n = 70_000
    X, _ = tsne.gaussian_mixture(n, dim=784, seed=4523)
    rp, col, val = tsne.joint_probabilities(X, perplexity=30.0, k=90, device="cuda")
    del X
    assert rp[-1] == col.size and abs(float(val.astype(np.float64).sum()) - 1.0) < 1e-4
    yx, yy = tsne.initial_embedding(n, seed=4523, std=1.0)
    theta = 0.5
    t = tsne.TSNE(rp, col, val, (yx, yy), min_theta=theta, max_theta=theta)
    try:
        dx, dy, z = t.timeStep(return_derivative=True)
        y1 = t.embeddedPoints.astype(np.float64)
    finally:
        t.close()
    x0, y0 = yx.astype(np.float64), yy.astype(np.float64)
    odx, ody, onx, ony, oz = O.tsne_step(x0, y0, x0, y0, rp, col, val, theta, 16.0, n / 15.0,
                                         0.5)
    qt = O.QuadTree(x0, y0, 16)
    rx, ry, _ = qt.forces(O.KERNEL_TSNE, 1.0, 1.0, 0.0, theta, nthreads=8)[:3]
    ax, ay = O.tsne_attraction(x0, y0, rp, col, val, 16.0)
    scale = np.hypot(np.abs(4.0 / oz * rx) + np.abs(ax), np.abs(4.0 / oz * ry) + np.abs(ay))
    e = np.hypot(dx - odx, dy - ody) / scale
    assert np.median(e) < 1e-5 and np.quantile(e, 0.99) < 1e-4, (np.median(e), e.max())
    assert abs(z - oz) < 1e-6 * oz
    e = np.hypot(y1[:, 0] - onx, y1[:, 1] - ony) / ((n / 15.0) * scale + 1e-6 * np.hypot(x0, y0))
    assert np.median(e) < 1e-5 and np.quantile(e, 0.99) < 1e-4


def test_config4_256_systems_sampled_sinks(gpu):
    ps = batched_systems(256)
    sizes = [p.n for p in ps]
    x, y, m = (np.concatenate([getattr(p, f) for p in ps]) for f in ("x", "y", "m"))
    ctx = BHContext(len(x), kernel="gravity", softening=EPS, sizes=sizes)
    ctx.set_particles(x, y, m)
    ctx.build()
    theta = 0.5
    ax, ay, _ = ctx.forces(theta, want_total=False)
    ns, off = ctx.batch_info()
    assert ns == 256 and int(off[-1]) == len(x) > 10_000_000
    worst = 0.0
    export = batched_export(ctx)
    for k, p in enumerate(ps):
        s = sample_indices(p.n, 16, seed=k)
        wx, wy, bx, by = exact_walk_system(ctx, k, p.x, p.y, p.m, "gravity", theta, s, eps=EPS,
                                           export=export)
        lo = int(off[k])
        gx, gy = ax[lo + s].astype(np.float64), ay[lo + s].astype(np.float64)
        e = np.maximum(np.abs(gx - wx) / np.where(bx > 0, bx, 1.0),
                       np.abs(gy - wy) / np.where(by > 0, by, 1.0))
        worst = max(worst, float(e.max()))
    print(f"config 4: 256 systems x 16 sinks, worst error {worst:.2e} of |contribution| sums")
    assert worst <= WALK_TOL
    ctx.close()


def _coefficients_of_records(rec):
    """c_j (Re, Im interleaved, [n, 14]) from the walk's coefficient records (the inverse of the
    record layout bh_learned_export documents)."""
    r = rec.astype(np.float64)
    c = np.zeros((r.shape[0], 14))
    c[:, 0], c[:, 2] = (r[:, 0] + r[:, 2]) / 2, (r[:, 0] - r[:, 2]) / 2  # a1, a2
    c[:, 1], c[:, 3] = (r[:, 3] - r[:, 1]) / 2, (r[:, 3] + r[:, 1]) / 2  # b1, b2
    c[:, 4:8] = r[:, 4:8]  # a3 b3 a4 b4
    c[:, 8], c[:, 10] = (r[:, 8] + r[:, 10]) / 2, (r[:, 8] - r[:, 10]) / 2  # a5, a6
    c[:, 9], c[:, 11] = (r[:, 11] - r[:, 9]) / 2, (r[:, 11] + r[:, 9]) / 2  # b5, b6
    c[:, 12:14] = r[:, 12:14]  # a7 b7
    return c


def test_config5_learned_4m(gpu):
    from masterthesislucas_amd.learned import FarFieldMLP

    n, theta = 1 << 22, 0.9
    ctx = BHContext(n, kernel="gravity", tree_mode="quadtree", softening=EPS)
    ctx.learned_enable(True)
    tr = plummer(n, a=1.0, seed=SEED_PLUMMER + 1)
    ctx.set_particles(tr.x, tr.y, tr.m)
    ctx.build()
    net = FarFieldMLP(max_samples=(1 << 21) + 4096 * 1024, seed=SEED_PLUMMER)
    added, rnd = 0, 0
    while added < 2_000_000:
        sinks = sample_indices(n, 4096, seed=1000 + rnd).astype(np.uint32)
        added += net.collect(ctx, theta, sinks, max_pairs_per_sink=1024)
        rnd += 1
    loss = net.train(epochs=8, batch=4096, lr=2e-3, seed=7)
    del tr
    p = plummer(n, a=1.0, seed=SEED_PLUMMER)
    ctx.set_particles(p.x, p.y, p.m)
    ctx.build()
    s = sample_indices(n, 4096)
    dx, dy, _ = ctx.direct_fp64(s)
    lx, ly, _ = ctx.forces_learned(net, theta)
    pn, pp = ctx.counters()
    # exact: the oracle's learned walk over the GPU's node records and coefficient records
    _, rec = ctx.learned_export(net)
    _, perm = ctx.morton()
    g = ctx.tree()
    _, side = ctx.bounds()
    where = np.empty(n, np.int64)
    where[perm] = np.arange(n)
    tree = dict(level=g["level"], first=g["first"], count=g["count"], skip=g["skip"],
                pp=((g["flags"] >> 1) & 1).astype(np.uint8), cx=g["cx"], cy=g["cy"],
                mass=g["mass"])
    coef = _coefficients_of_records(rec)
    ox, oy, cnt = O.bh_traverse_learned(p.x[perm], p.y[perm], p.m[perm], tree, side,
                                        O.inv_theta2_f32(theta), coef, eps2=EPS * EPS,
                                        sinks=where[s])
    w = exact_walk(ctx, p.x, p.y, p.m, "gravity", theta, eps=EPS, sinks=s)
    assert int(cnt[2]) == int(cnt[0]) == w["pn"]
    # each correction is at most theta^2 sum_j |c_j| times its monopole
    cmax = float(np.max(np.abs(coef).sum(1)))
    ex = np.abs(lx[s] - ox) / w["abs_x"]
    ey = np.abs(ly[s] - oy) / w["abs_y"]
    bar = WALK_TOL * (1.0 + theta * theta * cmax)
    assert max(ex.max(), ey.max()) <= bar, (ex.max(), ey.max(), bar)
    learned = float(np.median(rel_err(lx[s], ly[s], dx, dy)))
    ax, ay, _ = ctx.forces(theta, want_total=False)
    plain = float(np.median(rel_err(ax[s], ay[s], dx, dy)))
    bx, by, _ = ctx.forces(0.5, want_total=False)
    plain05 = float(np.median(rel_err(bx[s], by[s], dx, dy)))
    print(f"config 5: {added} pairs, loss {loss:.3e}; median error learned(0.9) {learned:.3e}, "
          f"plain(0.9) {plain:.3e}, plain(0.5) {plain05:.3e}; walk bar {bar:.2e} "
          f"(worst {max(ex.max(), ey.max()):.2e})")
    assert learned < plain / 3 and learned < plain05
    net.close()
    ctx.close()
