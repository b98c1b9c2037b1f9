"""Learned far-field correction on the GPU (include/bh.h, bh_mlp_* / bh_learned_*).

Every stage is checked against the FP64 oracle (oracle/bh_oracle.c), whose learned functions
are pinned in tests/test_oracle.py (basis and loss gradients by finite differences, the
quadrupole coefficients against brute-force cell sums):
  - cell features of the GPU's tree against orc_cell_features on the same node records;
  - the network (FP32 kernels and the BF16 MFMA inference) against orc_mlp_forward, the
    coefficient records the walk reads against the FP64 forward of the exported features;
  - the loss gradient against orc_learned_loss_grad, Adam against orc_adam;
  - the data-set walk and every label against orc_pair_label on the pairs the single-precision
    walk accepts;
  - the corrected walk against orc_bh_traverse_learned over the GPU's node records with the
    same coefficients (equal counts, per-particle error within the exact-walk bar);
  - and what a trained network gains over the plain monopole walk at the same theta.
"""
import numpy as np
import pytest

from masterthesislucas_amd import BHContext
from masterthesislucas_amd.learned import N_OUT, FarFieldMLP, flatten, unflatten
from masterthesislucas_amd.workloads import plummer, sample_indices
from oracle import oracle as O
from tests.helpers import WALK_TOL, exact_walk, rel_err

pytestmark = pytest.mark.gpu

EPS = float(np.float32(1e-3))


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = {"W1": rng.normal(0, 0.4, (10, 64)), "b1": rng.normal(0, 0.1, 64),
         "W2": rng.normal(0, 0.15, (64, 64)), "b2": rng.normal(0, 0.1, 64),
         "W3": rng.normal(0, 0.15, (64, N_OUT)) * scale,
         "b3": rng.normal(0, 0.1, N_OUT) * scale}
    return {k: v.astype(np.float32) for k, v in p.items()}


def _f64(p):
    return {k: v.astype(np.float64) for k, v in p.items()}


def _X(n, seed, cols=10):
    rng = np.random.default_rng(seed)
    return rng.normal(0, 1, (n, cols)).astype(np.float32)


def _records(y):
    """The walk's coefficient record of network outputs y [n, 14] (bh_learned_export)."""
    y = np.atleast_2d(y)
    r = np.zeros((y.shape[0], 16), y.dtype)
    r[:, 0], r[:, 1] = y[:, 0] + y[:, 2], y[:, 3] - y[:, 1]
    r[:, 2], r[:, 3] = y[:, 0] - y[:, 2], y[:, 1] + y[:, 3]
    r[:, 4:8] = y[:, 4:8]
    r[:, 8], r[:, 9] = y[:, 8] + y[:, 10], y[:, 11] - y[:, 9]
    r[:, 10], r[:, 11] = y[:, 8] - y[:, 10], y[:, 9] + y[:, 11]
    r[:, 12:14] = y[:, 12:14]
    return r


def test_mlp_forward_fp32_matches_oracle(gpu):
    net = FarFieldMLP(max_samples=16)
    p = _params(1)
    net.set_params(flatten(p))
    X = _X(1000, 2)
    y = net.forward(X)
    assert y.shape == (1000, 14)
    ref = O.mlp_forward(X.astype(np.float64), _f64(p))
    assert np.max(np.abs(y - ref) / (1.0 + np.abs(ref))) < 1e-5
    net.close()


def test_mlp_forward_mfma_bf16_within_bf16_bounds(gpu):
    net = FarFieldMLP(max_samples=16)
    p = _params(3)
    net.set_params(flatten(p))
    X = _X(1001, 4)  # not a multiple of 64: the tail rows are padding
    y = net.forward(X, mfma=True)
    ref = O.mlp_forward(X.astype(np.float64), _f64(p))
    err = np.abs(y - ref)
    scale = np.abs(ref).max()
    # BF16 inputs, weights and activations (8-bit mantissa), FP32 accumulation
    assert np.median(err) < 1e-2 * scale and err.max() < 6e-2 * scale, (np.median(err), err.max(), scale)
    assert np.corrcoef(y.ravel(), ref.ravel())[0, 1] > 0.999
    net.close()


@pytest.mark.parametrize("terms", [3, 7])
def test_learned_loss_gradients_match_oracle(gpu, terms):
    net = FarFieldMLP(max_samples=16)
    net.terms = terms
    assert net.terms == terms
    p = _params(5)
    net.set_params(flatten(p))
    X = _X(3000, 6, cols=13)
    w = np.random.default_rng(6).normal(0, 1, (3000, 2))
    X[:, 10:12] = (w / np.hypot(w[:, :1], w[:, 1:])).astype(np.float32)  # unit directions
    X[:, 12] = np.random.default_rng(7).uniform(0.05, 1.0, 3000).astype(np.float32)
    T = np.random.default_rng(8).normal(0, 0.5, (3000, 2)).astype(np.float32)
    loss, g = net.loss_grad(X, T)
    ref_loss, ref_g = O.learned_loss_grad(X.astype(np.float64), T.astype(np.float64), _f64(p),
                                          terms=terms)
    assert abs(loss - ref_loss) < 1e-5 * ref_loss
    gg = unflatten(g)
    for k in ref_g:
        num = np.linalg.norm(gg[k].astype(np.float64) - ref_g[k])
        assert num < 1e-4 * np.linalg.norm(ref_g[k]) + 1e-9, k
    net.close()


def test_adam_matches_oracle(gpu):
    net = FarFieldMLP(max_samples=16)
    p0 = flatten(_params(8))
    assert p0.size == net.count == 5774
    net.set_params(p0)
    rng = np.random.default_rng(9)
    ref = p0.astype(np.float64)
    m1, m2 = np.zeros_like(ref), np.zeros_like(ref)
    for t in (1, 2, 3):
        g = rng.normal(0, 1, p0.size).astype(np.float32)
        net.adam_step(g, 1e-2)
        O.adam(ref, g.astype(np.float64), m1, m2, 1e-2, 0.9, 0.999, 1e-8, t)
    assert np.max(np.abs(net.get_params() - ref)) < 1e-5
    net.close()


def test_training_reduces_loss(gpu):
    """Samples whose labels are the basis expansion of coefficients that depend smoothly on the
    cell features: training must cut the loss tenfold, and the loss it reports is the data-set
    loss of the trained network."""
    rng = np.random.default_rng(10)
    n = 20000
    X = np.empty((n, 13), np.float32)
    X[:, :10] = rng.uniform(-1, 1, (n, 10))
    a = rng.uniform(0, 2 * np.pi, n)
    X[:, 10], X[:, 11] = np.cos(a), np.sin(a)
    X[:, 12] = rng.uniform(0.1, 1.0, n)
    c = np.zeros((n, 14))
    c[:, 0] = 0.75 * (X[:, 0] + X[:, 2])
    c[:, 2], c[:, 3] = 0.375 * (X[:, 0] - X[:, 2]), 0.75 * X[:, 1]
    c[:, 6] = np.sin(2 * X[:, 3]) * X[:, 5]
    T = np.empty((n, 2), np.float32)
    for i in range(n):
        b = O.learned_basis(X[i, 10], X[i, 11], X[i, 12])
        T[i, 0] = np.sum(c[i, 0::2] * b[0::2] - c[i, 1::2] * b[1::2])
        T[i, 1] = np.sum(c[i, 0::2] * b[1::2] + c[i, 1::2] * b[0::2])
    net = FarFieldMLP(max_samples=n, seed=11)
    net.set_dataset(X, T)
    l0, _ = net.loss_grad(X, T)
    loss = net.train(epochs=30, batch=1024, lr=3e-3, seed=12)
    assert loss < l0 / 10, (l0, loss)
    l1, _ = net.loss_grad(X, T)
    assert abs(loss - l1) < 1e-3 * max(loss, 1e-6) + 1e-7
    Xd, Td = net.dataset()
    assert np.array_equal(Xd, X) and np.array_equal(Td, T)
    net.close()


def _plummer_ctx(n, seed, a=1.0):
    p = plummer(n, a=a, seed=seed)
    ctx = BHContext(n, kernel="gravity", tree_mode="quadtree", softening=EPS)
    ctx.set_particles(p.x, p.y, p.m)
    ctx.learned_enable(True)
    ctx.build()
    return ctx, p


def _gpu_tree(ctx, p):
    """The GPU's node records and sorted particles, in the oracle's tree-dict form."""
    codes, perm = ctx.morton()
    g = ctx.tree()
    corner, side = ctx.bounds()
    tree = dict(level=g["level"], first=g["first"], count=g["count"], skip=g["skip"],
                pp=((g["flags"] >> 1) & 1).astype(np.uint8), cx=g["cx"], cy=g["cy"],
                mass=g["mass"])
    return dict(codes=codes, perm=perm, tree=tree, corner=corner, side=side,
                xs=p.x[perm], ys=p.y[perm], ms=p.m[perm])


def test_cell_features_match_oracle(gpu):
    n = 60000
    ctx, p = _plummer_ctx(n, seed=31)
    t = _gpu_tree(ctx, p)
    f, _ = ctx.learned_export()
    tr = t["tree"]
    mom = dict(cx=tr["cx"], cy=tr["cy"], mass=tr["mass"])
    ref = O.cell_features(t["xs"], t["ys"], t["ms"], tr, t["codes"], 32, t["corner"][0],
                          t["corner"][1], t["side"], mom)
    assert f.shape == ref.shape and np.all(np.isfinite(f))
    # second moments: the upward pass in FP32; third moments: FP32 sums about the FP32 centre
    err = np.abs(f - ref)
    assert np.all(err[:, :3] <= 2e-5 + 1e-4 * np.abs(ref[:, :3])), float(err[:, :3].max())
    assert np.all(err[:, 3:7] <= 2e-5 + 1e-4 * np.abs(ref[:, 3:7])), float(err[:, 3:7].max())
    assert np.all(err[:, 7:9] <= 1e-6), float(err[:, 7:9].max())  # |offset| <= 0.5
    assert np.array_equal(f[:, 9], ref[:, 9].astype(np.float32))
    # closed forms: a cell's centre of mass lies in its box, and S is positive semi-definite
    assert np.all(np.abs(ref[:, 7:9]) <= 0.5 + 1e-12)
    assert np.all(f[:, 0] >= 0) and np.all(f[:, 2] >= 0)
    assert np.all(f[:, 0] * f[:, 2] - f[:, 1] ** 2 >= -1e-6)
    ctx.close()


def test_cell_coefficients_match_network(gpu):
    ctx, p = _plummer_ctx(60000, seed=32)
    net = FarFieldMLP(max_samples=16)
    prm = _params(33)
    net.set_params(flatten(prm))
    f, rec = ctx.learned_export(net)
    ref = _records(O.mlp_forward(f.astype(np.float64), _f64(prm)))[:, :14]
    scale = np.abs(ref).max()
    # the record's last two words: the cell's M l^2 and side l as the walk takes them (the FP32
    # root side scaled by 2^-level; M l^2 rounded as (M l) l)
    g = ctx.tree()
    _, side = ctx.bounds()
    l = np.ldexp(np.float32(side), -g["level"]).astype(np.float32)
    assert np.array_equal(rec[:, 15], l)
    assert np.array_equal(rec[:, 14], (g["mass"] * l).astype(np.float32) * l)
    rec = rec[:, :14]
    err = np.abs(rec - ref)
    # BF16 MFMA inference against the FP64 forward of the same features
    assert np.median(err) < 1e-2 * scale and err.max() < 6e-2 * scale, (np.median(err), err.max())
    # and the same network evaluated through the FP32 kernels on the exported features agrees
    y32 = net.forward(f)
    assert np.max(np.abs(_records(y32)[:, :14] - ref)) < 1e-4 * scale
    net.close()
    ctx.close()


def _fp32_walk_pairs(tr, side, px, py, theta, band):
    """(sink, cell) pairs the collecting walk takes for one sink: the single-precision
    acceptance d2 > l2 / theta^2 of the traversal, in pre-order."""
    f32 = np.float32
    inv_t2 = O.inv_theta2_f32(theta)
    inv_b2 = f32(0.0) if band == 0 else f32(1.0 / (band * band))
    out, k, nn = [], 0, tr["first"].size
    while k < nn:
        sl = np.ldexp(f32(side), -int(tr["level"][k])).astype(f32)
        l2 = f32(sl * sl)
        dx, dy = f32(tr["cx"][k]) - f32(px), f32(tr["cy"][k]) - f32(py)
        d2 = f32(float(dx) * float(dx) + float(f32(dy * dy)))
        nxt = max(int(tr["skip"][k]), k + 1)
        if d2 > f32(l2 * inv_t2):
            if inv_b2 == 0 or not d2 > f32(l2 * inv_b2):
                out.append(k)
            k = nxt
        elif tr["pp"][k]:
            k = nxt
        else:
            k += 1
    return out


def test_collect_labels_match_oracle(gpu):
    n = 20000
    ctx, p = _plummer_ctx(n, seed=22)
    t = _gpu_tree(ctx, p)
    tr = t["tree"]
    where = np.empty(n, np.int64)
    where[t["perm"]] = np.arange(n)
    f, _ = ctx.learned_export()
    sinks = sample_indices(n, 300).astype(np.uint32)
    net = FarFieldMLP(max_samples=1 << 20)
    theta = 0.9
    added = net.collect(ctx, theta, sinks, band=0.0, max_pairs_per_sink=4096)
    w = exact_walk(ctx, p.x, p.y, p.m, "gravity", theta, eps=EPS, sinks=sinks)
    assert added == w["pn"]  # one pair per accepted particle-node interaction
    X, T = net.dataset()
    assert X.shape == (added, 13) and T.shape == (added, 2)
    # the first 12 sinks pair by pair: the cell's features, offset features and FP64 label
    row = 0
    side32 = float(np.float32(t["side"]))
    for s in sinks[:12]:
        i = int(where[s])
        px, py = float(t["xs"][i]), float(t["ys"][i])
        cells = _fp32_walk_pairs(tr, side32, px, py, theta, 0.0)
        for k in cells:
            lk = np.ldexp(side32, -int(tr["level"][k]))
            ref = O.pair_label(t["xs"], t["ys"], t["ms"], tr["first"][k], tr["count"][k],
                               float(tr["cx"][k]), float(tr["cy"][k]), float(tr["mass"][k]), lk,
                               px, py, self_index=i, eps2=EPS * EPS)
            assert np.array_equal(X[row, :10], f[k]), (s, k)
            assert np.allclose(X[row, 10:13], ref[2:5], rtol=1e-6, atol=1e-7), (s, k)
            # FP64 label rounded to FP32; in these units the monopole itself is 1 / q^2, so
            # the FP64 sums' rounding is far below 1e-9 / q^2
            assert np.allclose(T[row], ref[:2], rtol=1e-5, atol=1e-9 / ref[4] ** 2), \
                (s, k, T[row], ref[:2])
            row += 1
    # with a band only pairs with l/r >= band remain
    net.clear()
    in_band = net.collect(ctx, theta, sinks, band=0.7, max_pairs_per_sink=4096)
    assert 0 < in_band < added
    X, T = net.dataset()
    assert np.all(X[:, 12] >= 0.7 * (1 - 1e-6)) and np.all(X[:, 12] < 0.9 * (1 + 1e-6))
    assert np.all(np.isfinite(T))
    net.close()
    ctx.close()


@pytest.mark.parametrize("variant,band,terms", [(0, 0.0, 7), (4, 0.0, 7), (4, 0.5, 7),
                                                (4, 0.0, 3), (8, 0.5, 3)])
def test_learned_walk_matches_oracle(gpu, variant, band, terms):
    """A network whose last layer is zero outputs its bias for every cell: the corrected walk
    with those constant coefficients against orc_bh_traverse_learned on the GPU's tree (with the
    coefficients of unused terms zero for a three-term network)."""
    n, theta = 30000, 0.9
    ctx, p = _plummer_ctx(n, seed=34)
    ctx.set_variant(variant)
    t = _gpu_tree(ctx, p)
    net = FarFieldMLP(max_samples=16)
    net.terms = terms
    prm = _params(35)
    prm["W3"][:] = 0.0
    prm["b3"][:] = np.random.default_rng(36).uniform(-1, 1, N_OUT).astype(np.float32)
    net.set_params(flatten(prm))
    _, rec = ctx.learned_export(net)
    assert np.array_equal(rec[:, :14],
                          np.broadcast_to(_records(prm["b3"][None, :])[0, :14], (rec.shape[0], 14)))
    ax, ay, _ = ctx.forces_learned(net, theta, band=band)
    pn, pp = ctx.counters()
    c = prm["b3"].astype(np.float64)
    c[2 * terms:] = 0.0
    coef = np.broadcast_to(c, (t["tree"]["first"].size, N_OUT))
    inv_b2 = 0.0 if band == 0 else 1.0 / (band * band)
    bx, by, cnt = O.bh_traverse_learned(t["xs"], t["ys"], t["ms"], t["tree"], t["side"],
                                        O.inv_theta2_f32(theta), coef, inv_band2=inv_b2,
                                        eps2=EPS * EPS)
    w = exact_walk(ctx, p.x, p.y, p.m, "gravity", theta, eps=EPS)
    assert (int(pn), int(pp)) == (int(cnt[0]), int(cnt[1])) == (w["pn"], w["pp"])
    assert 0 < int(cnt[2]) <= int(cnt[0]) and (band > 0) == (int(cnt[2]) < int(cnt[0]))
    rx, ry = np.empty(n), np.empty(n)
    rx[t["perm"]], ry[t["perm"]] = bx, by
    # FP32 correction terms: each at most (l/r)^2 |c| times its monopole, so the bar of the
    # exact walk scaled by the coefficients' size holds
    cmax = 4.0 * np.abs(prm["b3"]).sum()
    ex = np.abs(ax - rx) / w["abs_x"]
    ey = np.abs(ay - ry) / w["abs_y"]
    assert max(ex.max(), ey.max()) <= WALK_TOL * (1.0 + cmax), (ex.max(), ey.max())
    # and the correction is really there: the plain walk differs by far more than that bar
    px_, py_, _ = ctx.forces(theta, want_total=False)
    assert np.median(np.abs(px_ - rx) / w["abs_x"]) > 100 * WALK_TOL
    net.close()
    ctx.close()


@pytest.mark.parametrize("variant", [0, 4])
def test_zero_network_reproduces_plain_traversal(gpu, variant):
    ctx, _ = _plummer_ctx(60000, seed=21)
    ctx.set_variant(variant)
    net = FarFieldMLP(max_samples=16)
    p = _params(13)
    p["W3"][:] = 0.0
    p["b3"][:] = 0.0
    net.set_params(flatten(p))
    ax, ay, _ = ctx.forces(0.9, want_total=False)
    c_plain = ctx.counters()
    lx, ly, _ = ctx.forces_learned(net, 0.9)
    assert ctx.counters() == c_plain
    assert np.array_equal(ax, lx) and np.array_equal(ay, ly)
    net.close()
    ctx.close()


def test_learned_requires_quadtree_gravity(gpu):
    ctx = BHContext(1024, kernel="gravity", tree_mode="morton_fixed", softening=EPS)
    with pytest.raises(Exception):
        ctx.learned_enable(True)
    ctx.close()


@pytest.mark.parametrize("terms", [7, 3])
def test_trained_network_beats_plain_monopole(gpu, terms):
    """Trained on one Plummer realisation, evaluated on another: at theta = 0.9 the learned walk
    must cut the plain walk's median error against the FP64 direct sum at least threefold."""
    theta = 0.9
    ctx, _ = _plummer_ctx(150000, seed=23)
    net = FarFieldMLP(max_samples=1 << 21, seed=24)
    net.terms = terms
    added = net.collect(ctx, theta, sample_indices(150000, 6000, seed=25).astype(np.uint32))
    assert added > 100000
    loss = net.train(epochs=12, batch=4096, lr=2e-3, seed=26)
    assert np.isfinite(loss)
    ctx.close()
    ctx, _ = _plummer_ctx(150000, seed=27)
    sample = sample_indices(150000, 2048, seed=28).astype(np.uint32)
    dx, dy, _ = ctx.direct_fp64(sample)

    def err(ax, ay):
        return float(np.median(rel_err(ax[sample], ay[sample], dx, dy)))

    plain = err(*ctx.forces(theta, want_total=False)[:2])
    plain05 = err(*ctx.forces(0.5, want_total=False)[:2])
    learned = err(*ctx.forces_learned(net, theta)[:2])
    print(f"[learned] terms={terms} theta={theta} pairs={added} loss={loss:.3e} plain={plain:.3e} "
          f"learned={learned:.3e} plain(0.5)={plain05:.3e}", flush=True)
    assert learned < plain / 3, (plain, learned)
    net.close()
    ctx.close()
