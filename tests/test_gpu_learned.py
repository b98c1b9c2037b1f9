"""Learned far-field correction on the GPU (include/bh.h, bh_mlp_* / bh_learned_*).

The network kernels are checked against the FP64 oracle (oracle/bh_oracle.c: orc_mlp_forward,
orc_mlp_loss_grad, orc_adam), whose own gradients are pinned by finite differences in
tests/test_oracle.py. The data-set walk is checked against the oracle's interaction counts, the
learned traversal against the plain one (a network that outputs zero must reproduce it bit for
bit), and a trained network against the plain method at the same theta (the gain).
"""
import numpy as np
import pytest

from masterthesislucas_amd import BHContext
from masterthesislucas_amd.learned import FarFieldMLP, flatten, unflatten
from masterthesislucas_amd.workloads import plummer, sample_indices
from oracle import oracle as O
from tests.helpers import oracle_tree

pytestmark = pytest.mark.gpu

EPS = float(np.float32(1e-3))


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = {"W1": rng.normal(0, 0.4, (10, 64)), "b1": rng.normal(0, 0.1, 64),
         "W2": rng.normal(0, 0.15, (64, 64)), "b2": rng.normal(0, 0.1, 64),
         "W3": rng.normal(0, 0.15, (64, 2)) * scale, "b3": rng.normal(0, 0.1, 2) * scale}
    return {k: v.astype(np.float32) for k, v in p.items()}


def _X(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(0, 1, (n, 10)).astype(np.float32)


def test_mlp_forward_fp32_matches_oracle(gpu):
    net = FarFieldMLP(max_samples=16)
    p = _params(1)
    net.set_params(flatten(p))
    X = _X(1000, 2)
    y = net.forward(X)
    ref = O.mlp_forward(X.astype(np.float64), {k: v.astype(np.float64) for k, v in p.items()})
    assert np.max(np.abs(y - ref) / (1.0 + np.abs(ref))) < 1e-5
    net.close()


def test_mlp_forward_mfma_bf16_within_bf16_bounds(gpu):
    net = FarFieldMLP(max_samples=16)
    p = _params(3)
    net.set_params(flatten(p))
    X = _X(1001, 4)  # not a multiple of 64: the tail block is padded
    y = net.forward(X, mfma=True)
    ref = O.mlp_forward(X.astype(np.float64), {k: v.astype(np.float64) for k, v in p.items()})
    err = np.abs(y - ref)
    scale = np.abs(ref).max()
    # BF16 inputs, weights and activations (8-bit mantissa), FP32 accumulation
    assert np.median(err) < 1e-2 * scale and err.max() < 6e-2 * scale, (np.median(err), err.max(), scale)
    assert np.corrcoef(y.ravel(), ref.ravel())[0, 1] > 0.999
    net.close()


def test_mlp_gradients_match_oracle(gpu):
    net = FarFieldMLP(max_samples=16)
    p = _params(5)
    net.set_params(flatten(p))
    X = _X(3000, 6)
    T = np.random.default_rng(7).normal(0, 0.5, (3000, 2)).astype(np.float32)
    loss, g = net.loss_grad(X, T)
    ref_loss, ref_g = O.mlp_loss_grad(X.astype(np.float64), T.astype(np.float64),
                                      {k: v.astype(np.float64) for k, v in p.items()})
    assert abs(loss - ref_loss) < 1e-5 * ref_loss
    gg = unflatten(g)
    for k in ref_g:
        num = np.linalg.norm(gg[k].astype(np.float64) - ref_g[k])
        assert num < 1e-4 * np.linalg.norm(ref_g[k]) + 1e-9, k
    net.close()


def test_adam_matches_oracle(gpu):
    net = FarFieldMLP(max_samples=16)
    p0 = flatten(_params(8))
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
    rng = np.random.default_rng(10)
    X = rng.uniform(-1, 1, (20000, 10)).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) * X[:, 2], X[:, 1] * X[:, 3] - 0.5 * X[:, 4] ** 2], 1)
    net = FarFieldMLP(max_samples=20000, seed=11)
    net.set_dataset(X, T.astype(np.float32))
    _, g0 = net.loss_grad(X, T)
    l0 = float(np.mean((net.forward(X) - T) ** 2))
    loss = net.train(epochs=30, batch=1024, lr=3e-3, seed=12)
    assert loss < l0 / 10, (l0, loss)
    # the reported loss is the data-set MSE of the trained network
    assert abs(loss - float(np.mean((net.forward(X) - T) ** 2))) < 1e-3 * max(loss, 1e-6) + 1e-7
    net.close()


def _plummer_ctx(n, seed, a=1.0):
    p = plummer(n, a=a, seed=seed)
    ctx = BHContext(n, kernel="gravity", tree_mode="quadtree", softening=EPS)
    ctx.set_particles(p.x, p.y, p.m)
    ctx.learned_enable(True)
    ctx.build()
    return ctx, p


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


def test_collect_walk_matches_oracle_counts(gpu):
    ctx, p = _plummer_ctx(20000, seed=22)
    sinks = sample_indices(20000, 300).astype(np.uint32)
    net = FarFieldMLP(max_samples=1 << 20)
    added = net.collect(ctx, 0.9, sinks, band=0.0, max_pairs_per_sink=4096)
    t = oracle_tree(p.x, p.y, p.m)
    sorted_pos = np.empty(20000, np.int64)
    sorted_pos[t["perm"]] = np.arange(20000)
    _, _, _, cnt = O.bh_traverse(t["xs"], t["ys"], t["ms"], t["tree"], t["mom"], t["side"],
                                 O.SEM_QUADTREE, O.KERNEL_GRAVITY, eps2=EPS * EPS, theta=0.9,
                                 sinks=sorted_pos[sinks], pp=t["pp"])
    assert added == int(cnt[0])  # one pair per accepted particle-node interaction
    # with a band only pairs with l/r >= band remain
    net.clear()
    in_band = net.collect(ctx, 0.9, sinks, band=0.7, max_pairs_per_sink=4096)
    assert 0 < in_band < added
    X, T = net.dataset()
    assert np.all(X[:, 2] >= 0.7 * (1 - 1e-6)) and np.all(X[:, 2] < 0.9 * (1 + 1e-6))
    assert np.all(np.isfinite(T))
    net.close()
    ctx.close()


def test_trained_network_beats_plain_monopole(gpu):
    theta = 0.9
    ctx, _ = _plummer_ctx(150000, seed=23)
    net = FarFieldMLP(max_samples=1 << 21, seed=24)
    added = net.collect(ctx, theta, sample_indices(150000, 6000, seed=25).astype(np.uint32))
    assert added > 100000
    loss = net.train(epochs=12, batch=4096, lr=2e-3, seed=26)
    assert np.isfinite(loss)
    ctx.close()
    # evaluate on a different realisation
    ctx, _ = _plummer_ctx(150000, seed=27)
    sample = sample_indices(150000, 2048, seed=28).astype(np.uint32)
    dx, dy, _ = ctx.direct_fp64(sample)
    dn = np.hypot(dx, dy)

    def err(ax, ay):
        return np.median(np.hypot(ax[sample] - dx, ay[sample] - dy) / dn)

    plain = err(*ctx.forces(theta, want_total=False)[:2])
    learned = err(*ctx.forces_learned(net, theta)[:2])
    print(f"[learned] theta={theta} pairs={added} loss={loss:.3e} plain={plain:.3e} "
          f"learned={learned:.3e}", flush=True)
    assert learned < 0.5 * plain, (plain, learned)
    net.close()
    ctx.close()
