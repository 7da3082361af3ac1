"""GPU exact kNN (the NeighborList producer) against the oracle's brute force.

Indices bit-exact and distances equal (double distances identical, float
distances the rounding of the oracle's double), self excluded, rows in
increasing distance with ties to the smaller index; across dimension counts
that cross the 64-coordinate staging boundary, k in every top-k capacity
class, float32 and float64 input, duplicate points (ties) and the resident
graph built from data.
"""
import numpy as np
import pytest

import oracle
from _data import normal_data

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,dim", [(300, 3), (1000, 10), (700, 50), (500, 70), (2000, 5)])
@pytest.mark.parametrize("k", [5, 15, 30, 40])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_knn_matches_oracle(umx, n, dim, k, dtype):
    x = normal_data(n * 7 + dim, n, dim).astype(dtype)
    idx, dist = umx.find_knn(x, k)
    ridx, rdist = oracle.knn_bruteforce(x.astype(np.float64), k)
    assert np.array_equal(idx, ridx)
    if dtype == np.float64:
        assert np.array_equal(dist, rdist)
    else:
        assert np.array_equal(dist, rdist.astype(np.float32))
    assert np.all(idx != np.arange(n)[:, None])


@pytest.mark.parametrize("case", ["normal", "grid_ties", "offset", "duplicates"])
@pytest.mark.parametrize("n,dim,k", [(512, 1, 16), (1001, 4, 15), (1500, 5, 1), (777, 63, 10), (2049, 64, 16),
                                     (3000, 50, 15)])
def test_knn_matrix_core_path(umx, case, n, dim, k):
    """float32 input with at most 64 coordinates, k <= 16 and n >= 512: the
    candidates come from the matrix cores (float32 dot products of centred
    points within a proven margin of the exact distances), exact distances
    rank them, and queries whose candidate lists cannot prove completeness
    (ties on a grid, duplicates, a large common offset that widens the
    margin) are searched again exactly.  Output bit-identical to the brute
    force in every case."""
    rng = np.random.default_rng(n * 131 + dim)
    if case == "normal":
        x = rng.normal(size=(n, dim))
    elif case == "grid_ties":
        x = rng.integers(0, 3, size=(n, dim)).astype(np.float64)
    elif case == "offset":
        x = rng.normal(size=(n, dim)) + 1000.0
    else:
        base = rng.normal(size=(n // 3 + 1, dim))
        x = base[rng.integers(0, base.shape[0], size=n)]
    x = x.astype(np.float32)
    idx, dist = umx.find_knn(x, k)
    ridx, rdist = oracle.knn_bruteforce(x.astype(np.float64), k)
    assert np.array_equal(idx, ridx)
    assert np.array_equal(dist, rdist.astype(np.float32))
    assert np.all(idx != np.arange(n)[:, None])


def test_knn_ties_go_to_smaller_index(umx):
    rng = np.random.default_rng(4)
    x = rng.integers(0, 3, size=(400, 3)).astype(np.float64)  # many duplicates and equal distances
    idx, dist = umx.find_knn(x, 20)
    ridx, rdist = oracle.knn_bruteforce(x, 20)
    assert np.array_equal(idx, ridx) and np.array_equal(dist, rdist)


def test_graph_from_data_equals_host_graph(umx):
    x = normal_data(77, 1500, 20)
    g1 = umx.Graph.from_data(x, 15)
    g1.neighbor_similarities()
    g1.combine(1.0)
    idx, dist = umx.find_knn(x, 15)
    g2 = umx.Graph(umx.NeighborList.from_knn(idx, dist))
    g2.neighbor_similarities()
    g2.combine(1.0)
    a, b = g1.export(), g2.export()
    assert np.array_equal(a.row_ptr, b.row_ptr) and np.array_equal(a.idx, b.idx) and np.array_equal(a.val, b.val)


def test_knn_argument_errors(umx):
    x = np.zeros((10, 2))
    with pytest.raises(umx.UmxError):
        umx.find_knn(x, 10)  # k must be < n
    with pytest.raises(umx.UmxError):
        umx.find_knn(np.zeros((100, 2)), 65)
