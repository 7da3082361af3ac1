"""Synthetic inputs shared by the tests (regenerated from seeds, never stored).

The reference's fixtures are regenerated from their recipes:
  * neighbor_similarities.cpp:8-23  N(0,1) in 5 dims, mt19937_64(nobs*k + conn*10)
  * combine_neighbor_sets.cpp:16-34 N(0,1) in 5 dims, mt19937_64(nobs*k), weights exp(-dist)
  * optimize_layout.cpp:16-56       N(0,1) in 5 dims, mt19937_64(nobs*k), similarities + union
  * README.md:27-34 (config C1)     2000 x 10 N(0,1), mt19937_64(1000)
libstdc++'s normal_distribution is used through the oracle library, so the
draws are the ones a g++ build of the reference tests would see.  Exact kNN is
computed by brute force (it equals the VP-tree result on continuous data).
"""
import numpy as np

import oracle


def normal_data(seed, nobs, ndim):
    return oracle.gen_normal(seed, nobs * ndim).reshape(nobs, ndim)


def knn_lists(x, k):
    idx, dist = oracle.knn_bruteforce(x, k)
    n = x.shape[0]
    rp = np.arange(n + 1, dtype=np.int64) * k
    return rp, idx.reshape(-1).copy(), dist.reshape(-1).copy()


def similarity_fixture(nobs, k, conn):
    seed = int(nobs * k + conn * 10)
    x = normal_data(seed, nobs, 5)
    return knn_lists(x, k)


def combine_fixture(nobs, k):
    x = normal_data(nobs * k, nobs, 5)
    rp, idx, dist = knn_lists(x, k)
    return rp, idx, np.exp(-dist)


def layout_fixture(nobs, k):
    """Symmetrised fuzzy graph of the reference's OptimizeTest (and its data)."""
    x = normal_data(nobs * k, nobs, 5)
    rp, idx, dist = knn_lists(x, k)
    w, _, _ = oracle.neighbor_similarities(rp, dist)
    orp, oidx, oval = oracle.combine_neighbor_sets(rp, idx, w, 1.0)
    return x, orp, oidx, oval


def c1_knn():
    """Config C1: 2000 x 10 N(0,1) (mt19937_64 seed 1000), exact kNN k=15."""
    x = normal_data(1000, 2000, 10)
    return x, knn_lists(x, 15)


def mixture(seed, n, dim, ncomp, dtype=np.float32):
    """Gaussian mixture: uniform weights, means ~ N(0, 25 I), per-component
    per-dimension std ~ U(0.5, 2) (numpy Generator, seeded)."""
    rng = np.random.default_rng(seed)
    means = rng.normal(0.0, 5.0, size=(ncomp, dim))
    sds = rng.uniform(0.5, 2.0, size=(ncomp, dim))
    lab = rng.integers(0, ncomp, size=n)
    x = means[lab] + rng.normal(size=(n, dim)) * sds[lab]
    return x.astype(dtype), lab


def random_ragged_graph(seed, n, max_k, dtype=np.float64, hubs=0):
    """Directed graph with unique non-self neighbours per row, variable row
    lengths (including empty rows) and optional hub targets."""
    rng = np.random.default_rng(seed)
    rows_idx, rows_val = [], []
    hub_ids = rng.choice(n, size=hubs, replace=False) if hubs else np.array([], dtype=np.int64)
    for i in range(n):
        k = int(rng.integers(0, max_k + 1))
        cand = rng.choice(n - 1, size=min(k, n - 1), replace=False)
        cand = np.where(cand >= i, cand + 1, cand)
        if hubs and rng.random() < 0.5:
            h = hub_ids[rng.integers(0, hubs)]
            if h != i and h not in cand:
                cand = np.append(cand, h)
        d = np.sort(rng.exponential(1.0, size=cand.size))
        rows_idx.append(cand.astype(np.int32))
        rows_val.append(d.astype(dtype))
    rp = np.zeros(n + 1, dtype=np.int64)
    rp[1:] = np.cumsum([r.size for r in rows_idx])
    idx = np.concatenate(rows_idx) if rp[-1] else np.zeros(0, np.int32)
    val = np.concatenate(rows_val) if rp[-1] else np.zeros(0, dtype)
    return rp, idx.astype(np.int32), val.astype(dtype)
