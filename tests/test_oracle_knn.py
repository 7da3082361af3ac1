"""Pins the oracle's exact brute-force kNN (CPU only).

* Neighbour lists equal scipy's cKDTree (a k-d tree search, a different
  algorithm) on continuous data, distances within 1e-12 relative, rows sorted
  by distance, the query itself excluded.
* Ties resolve to the smaller index: on an integer grid (many equal
  distances, all exact in double) every row equals a sort of the exact
  (squared distance, index) pairs.
* A query range equals those rows of the full search.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import oracle


@pytest.mark.parametrize("n,dim,k", [(500, 3, 5), (800, 20, 15), (300, 50, 30)])
def test_matches_kdtree(n, dim, k):
    x = np.random.default_rng(n + dim).normal(size=(n, dim))
    idx, dist = oracle.knn_bruteforce(x, k, nthreads=4)
    d_ref, i_ref = cKDTree(x).query(x, k + 1)
    assert np.array_equal(i_ref[:, 0], np.arange(n))  # continuous data: self first, at distance 0
    assert np.array_equal(idx, i_ref[:, 1:])
    np.testing.assert_allclose(dist, d_ref[:, 1:], rtol=1e-12)
    assert np.all(np.diff(dist, axis=1) >= 0)


def test_ties_go_to_smaller_index():
    g = np.stack(np.meshgrid(np.arange(6), np.arange(6)), -1).reshape(-1, 2).astype(np.float64)
    n, k = g.shape[0], 8
    idx, dist = oracle.knn_bruteforce(g, k, nthreads=2)
    for i in range(n):
        d2 = ((g - g[i]) ** 2).sum(axis=1)  # small integers: exact
        want = sorted((d2[j], j) for j in range(n) if j != i)[:k]
        assert idx[i].tolist() == [j for _, j in want]
        assert np.array_equal(dist[i], np.sqrt([d for d, _ in want]))


def test_query_range_equals_full_rows():
    x = np.random.default_rng(7).normal(size=(700, 12))
    idx, dist = oracle.knn_bruteforce(x, 10, nthreads=3)
    for q0, q1 in [(0, 1), (100, 350), (650, 700), (5, 5)]:
        ri, rd = oracle.knn_bruteforce_rows(x, 10, q0, q1, nthreads=4)
        assert np.array_equal(ri, idx[q0:q1]) and np.array_equal(rd, dist[q0:q1])
    with pytest.raises(ValueError):
        oracle.knn_bruteforce_rows(x, 10, 600, 701)
