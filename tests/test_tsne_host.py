"""Host-side t-SNE logic: the oracle's affinities (kNN, perplexity bisection, symmetrisation)
pinned by a full sort, closed forms and a dense matrix, and the library's theta schedule against
the oracle. No GPU needed."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.optimize import brentq

from masterthesislucas_amd import _lib, tsne
from oracle import oracle as orc


def test_knn_exact_matches_brute_force():
    rng = np.random.default_rng(3)
    X = rng.normal(size=(300, 20)).astype(np.float32)
    idx, d2 = orc.knn_exact(X, 12, chunk=64)
    X64 = X.astype(np.float64)
    D = ((X64[:, None, :] - X64[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(D, np.inf)
    ref = np.argsort(D, axis=1, kind="stable")[:, :12]
    assert np.array_equal(idx, ref)
    assert np.allclose(d2, np.take_along_axis(D, ref, 1), rtol=1e-12, atol=0)
    # a subset of rows gives the same lists; duplicated points tie at distance 0, by index
    sub = np.array([5, 0, 299])
    si, sd = orc.knn_exact(X, 12, rows=sub)
    assert np.array_equal(si, idx[sub]) and np.array_equal(sd, d2[sub])
    Xd = np.concatenate([X, X[:3], X[:3]])
    di, dd = orc.knn_exact(Xd, 3, rows=[0, 300])
    assert di[0].tolist() == [300, 303, int(ref[0, 0])] and dd[0, 0] == dd[0, 1] == 0.0
    assert di[1].tolist()[:2] == [0, 303]


def test_perplexity_bisection_reaches_target():
    rng = np.random.default_rng(4)
    d2 = np.sort(rng.uniform(0.5, 40.0, size=(200, 90)), axis=1)
    P, beta = orc.conditional_probabilities(d2, 30.0)
    assert np.allclose(P.sum(1), 1.0, atol=1e-12)
    H = -(P * np.log(P)).sum(1)
    assert np.allclose(np.exp(H), 30.0, rtol=1e-4)
    assert (beta > 0).all()


def test_perplexity_two_point_closed_form():
    # two neighbours at squared distances 0 and a: p = 1/(1+e^{-beta a}), entropy of a Bernoulli
    a, perp = 2.0, 1.5
    target = math.log(perp)

    def h(b):
        p = 1.0 / (1.0 + math.exp(-b * a))
        return -(p * math.log(p) + (1 - p) * math.log(1 - p)) - target

    b_star = brentq(h, 1e-9, 10.0, xtol=1e-14)
    _, beta = orc.conditional_probabilities(np.array([[0.0, a]]), perp, tol=1e-12)
    assert abs(beta[0] - b_star) < 1e-6 * b_star


def test_symmetrize_matches_dense():
    rng = np.random.default_rng(5)
    n, k = 40, 6
    idx = np.stack([rng.choice(np.delete(np.arange(n), i), k, replace=False) for i in range(n)])
    Pc = rng.uniform(size=(n, k))
    Pc /= Pc.sum(1, keepdims=True)
    rp, col, val = orc.symmetrize(idx, Pc)
    dense = np.zeros((n, n))
    for i in range(n):
        dense[i, idx[i]] = Pc[i]
    ref = (dense + dense.T) / (2 * n)
    got = np.zeros((n, n))
    for r in range(n):
        cols = col[rp[r]:rp[r + 1]]
        assert (np.diff(cols.astype(np.int64)) > 0).all()  # sorted, unique (tsne.h:462-520)
        got[r, cols] = val[rp[r]:rp[r + 1]]
    assert np.allclose(got, ref, rtol=1e-6, atol=1e-9)
    assert np.allclose(got, got.T)
    assert abs(val.sum(dtype=np.float64) - 1.0) < 1e-6
    assert rp[-1] == col.size == np.count_nonzero(ref)


def test_joint_probabilities_mixture():
    X, labels = tsne.gaussian_mixture(600, dim=30, seed=7)
    assert labels.max() < 10
    rp, col, val = orc.joint_probabilities(X, perplexity=10.0, k=30)
    assert rp.size == 601 and rp[-1] == col.size == val.size
    assert (val > 0).all()
    assert abs(val.sum(dtype=np.float64) - 1.0) < 1e-5
    rows = np.repeat(np.arange(600), np.diff(rp))
    assert not (rows == col).any()


def test_theta_schedule_matches_oracle():
    L = _lib.load()
    p = _lib.TsneParams()
    assert L.bh_tsne_params_default(C.byref(p), 70000, 0.5) == _lib.BH_OK
    assert abs(p.learning_rate - 70000 / 15.0) < 1e-12
    assert (p.momentum_early, p.momentum_late, p.exag_early, p.exag_late) == (0.5, 0.8, 16.0, 4.0)
    p.theta_min, p.theta_max = 0.3, 1.2
    assert L.bh_tsne_theta(C.byref(p), 0) == 1.2
    for t in range(1, 1300, 7):
        assert L.bh_tsne_theta(C.byref(p), t) == pytest.approx(
            orc.tsne_theta(0.3, 1.2, t - 1), abs=1e-15)
    # closed form: the midpoint of the schedule and its end
    assert orc.tsne_theta(0.3, 1.2, 0) == pytest.approx(1.2)
    assert orc.tsne_theta(0.3, 1.2, 999) == pytest.approx(0.3)
    assert orc.tsne_theta(0.3, 1.2, 5000) == pytest.approx(0.3)
    # symmetric about the middle of the schedule: theta(t) + theta(999 - t) = min + max
    for t in (1, 100, 333, 499):
        assert orc.tsne_theta(0.3, 1.2, t) + orc.tsne_theta(0.3, 1.2, 999 - t) == \
            pytest.approx(1.5, abs=1e-14)


def test_nmae_closed_form():
    """TsneTest::getNMAE: mean over the points of |naive - approx| / |naive|."""
    naive = np.array([[1.0, 0.0], [0.0, 2.0], [3.0, 4.0]])
    approx = np.array([[1.0, 0.1], [0.0, 1.0], [3.0, 4.0]])
    assert tsne.nmae(naive, approx) == pytest.approx((0.1 / 1.0 + 1.0 / 2.0 + 0.0) / 3.0, abs=1e-15)
    assert tsne.nmae(naive, naive) == 0.0
    # scale invariance of each term, and the mean (not the max) over the points
    assert tsne.nmae(7.0 * naive, 7.0 * approx) == pytest.approx(tsne.nmae(naive, approx), abs=1e-15)
    with pytest.raises(ValueError):
        tsne.nmae(naive, approx[:2])
