"""t-SNE input affinities on the GPU (bh_tsne_affinities) against the oracle's restatement of
openTSNE's algorithm (oracle.knn_exact / conditional_probabilities / symmetrize, pinned in
tests/test_tsne_host.py).

- kNN: the sorted FP64 distances equal the oracle's to 1e-12 for every row, and the neighbour
  indices equal it wherever the distance is not tied (ties go to the lower index on both sides).
- Bandwidths: every row's entropy is within the bisection tolerance of log(perplexity), and beta
  equals the oracle's run on the GPU's distances (the same bisection) for all but near-boundary
  rows.
- P: row_ptr and columns equal the oracle's symmetrisation of the GPU's lists, values to FP32
  rounding; P sums to 1.
Config 3's size (70,000 points in 784-D, perplexity 30, k = 90) runs on 512 sampled rows for
the kNN and on every row for the rest.
"""
import math
import time

import numpy as np
import pytest

from masterthesislucas_amd import tsne
from masterthesislucas_amd._lib import BHError
from masterthesislucas_amd.workloads import sample_indices
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _check_knn(idx, d2, ridx, rd2):
    assert np.allclose(d2, rd2, rtol=1e-12, atol=0), float(np.max(np.abs(d2 - rd2) / rd2))
    tied = np.zeros(d2.shape, bool)
    near = np.abs(np.diff(rd2, axis=1)) <= 1e-12 * rd2[:, 1:]
    tied[:, 1:] |= near
    tied[:, :-1] |= near
    # the k-th distance may tie with the (k+1)-th outside the list: compare sets there
    assert np.array_equal(idx[~tied], ridx[~tied])


def _cond_p(d2, beta):
    dr = d2 - d2[:, :1]
    E = np.exp(-beta[:, None] * dr)
    return E / E.sum(1, keepdims=True)


def _check_bandwidths(d2, beta, perplexity, tol=1e-5):
    P = _cond_p(d2, beta)
    dr = d2 - d2[:, :1]
    H = np.log(np.exp(-beta[:, None] * dr).sum(1)) + beta * (dr * P).sum(1)
    return np.abs(H - math.log(perplexity))


def _check_csr(a, n, perplexity):
    Pc = _cond_p(a["d2"], a["beta"])
    rp, col, val = O.symmetrize(a["idx"].astype(np.int64), Pc)
    assert np.array_equal(a["row_ptr"], rp)
    assert np.array_equal(a["col"], col)
    assert np.allclose(a["val"], val, rtol=2e-7, atol=0)
    assert abs(a["val"].sum(dtype=np.float64) - 1.0) < 1e-5
    rows = np.repeat(np.arange(n), np.diff(a["row_ptr"].astype(np.int64)))
    assert not (rows == a["col"]).any()
    assert (np.diff(a["col"].astype(np.int64))[np.diff(rows) == 0] > 0).all()


@pytest.mark.parametrize("n,dim,perp,k", [(3000, 50, 20.0, 60), (1000, 33, 5.0, 15)])
def test_affinities_match_oracle(gpu, n, dim, perp, k):
    X, _ = tsne.gaussian_mixture(n, dim=dim, seed=41 + dim)
    a = tsne.affinities(X, perp, k)
    ridx, rd2 = O.knn_exact(X, k)
    _check_knn(a["idx"].astype(np.int64), a["d2"], ridx, rd2)
    err = _check_bandwidths(a["d2"], a["beta"], perp)
    assert err.max() < 1.0001e-5, float(err.max())
    _, rbeta = O.conditional_probabilities(a["d2"], perp)
    rel = np.abs(a["beta"] - rbeta) / rbeta
    assert np.median(rel) < 1e-12 and rel.max() < 1e-3
    _check_csr(a, n, perp)


def test_affinities_duplicates_and_ragged_sizes(gpu):
    """Groups of exact duplicates (distance-0 ties, broken by index), n and dim not multiples
    of the tile sizes, and a scale that needs the FP16 range guard (|x| up to 1e6)."""
    rng = np.random.default_rng(5)
    base = rng.normal(0, 1e5, (130, 77)).astype(np.float32)
    X = np.concatenate([base, base[:20], base[:20], base[5:9]]).astype(np.float32)
    n, k = X.shape[0], 7
    a = tsne.affinities(X, 3.0, k)
    ridx, rd2 = O.knn_exact(X, k)
    _check_knn(a["idx"].astype(np.int64), a["d2"], ridx, rd2)
    assert a["d2"][0, 0] == 0.0 and a["idx"][0, 0] == 130  # first duplicate of point 0
    # a row whose nearest distance is shared by at least `perplexity` neighbours cannot reach
    # the target entropy (it stays above log of that count): both sides double beta to the
    # iteration limit there; every other row converges
    ties = (a["d2"] == a["d2"][:, :1]).sum(1)
    err = _check_bandwidths(a["d2"], a["beta"], 3.0)
    assert err[ties < 3].max() < 1.0001e-5
    _, rbeta = O.conditional_probabilities(a["d2"], 3.0)
    assert (ties >= 4).any() and np.array_equal(a["beta"][ties >= 4], rbeta[ties >= 4])
    assert np.allclose(a["beta"], rbeta, rtol=1e-3, atol=0)
    _check_csr(a, n, 3.0)


def test_affinities_reject_bad_input(gpu):
    X = np.random.default_rng(6).normal(size=(50, 8)).astype(np.float32)
    with pytest.raises(BHError):
        tsne.affinities(X, 5.0, 50)  # k >= n
    with pytest.raises(BHError):
        tsne.affinities(np.zeros((50, 1100), np.float32), 5.0, 5)  # dim > 1024
    Y = X.copy()
    Y[3, 2] = np.nan
    with pytest.raises(BHError):
        tsne.affinities(Y, 5.0, 5)


def test_affinities_config3_70k_784(gpu):
    n, perp, k = 70_000, 30.0, 90
    X, _ = tsne.gaussian_mixture(n, dim=784, seed=4523)
    t0 = time.perf_counter()
    a = tsne.affinities(X, perp, k)
    t = time.perf_counter() - t0
    rows = sample_indices(n, 512, seed=3)
    ridx, rd2 = O.knn_exact(X, k, rows=rows)
    _check_knn(a["idx"][rows].astype(np.int64), a["d2"][rows], ridx, rd2)
    err = _check_bandwidths(a["d2"], a["beta"], perp)
    assert err.max() < 1.0001e-5
    _check_csr(a, n, perp)
    print(f"config 3 affinities: {t:.3f} s for n={n}, 784-D, k={k} (host copies included); "
          f"nnz {a['col'].size}")
