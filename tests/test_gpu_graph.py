"""GPU parity of the graph stages against the oracle, through the C ABI.

K1 similarities: rho, sigma and weights within 1e-6 relative of the oracle's
    double path (the GPU sums each row with a lane-parallel tree, the oracle
    sequentially; rho involves no sum and must match exactly).
K2 combine:      row pointers, indices, order and weights bit-exact.
K3 epochs:       EpochData bit-exact.
"""
import itertools

import numpy as np
import pytest

import oracle
from _data import c1_knn, combine_fixture, layout_fixture, random_ragged_graph, similarity_fixture

pytestmark = pytest.mark.gpu

REL = 1e-6  # similarities tolerance (relative)


def _gpu_sim(umx, rp, dist, conn=1.0, bw=1.0, mkds=1e-3, newton=True):
    nl = umx.NeighborList(rp, np.zeros(dist.size, np.int32), dist.copy())
    o = umx.NeighborSimilaritiesOptions(local_connectivity=conn, bandwidth=bw, min_k_dist_scale=mkds)
    rho, sigma = umx.neighbor_similarities(nl, o, use_newton=newton, return_rho_sigma=True)
    return nl.val, rho, sigma


def _close(a, b, rel):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    ok = both_nan | (np.abs(a - b) <= rel * np.maximum(np.abs(b), 1e-300))
    assert ok.all(), f"max rel err {np.nanmax(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))}"


@pytest.mark.parametrize("nobs,k,conn", list(itertools.product([50, 100, 200], [5, 10, 20],
                                                               [0.4, 1.0, 1.3, 2.0, 2.5])))
@pytest.mark.parametrize("newton", [True, False])
def test_similarities_reference_grid(umx, nobs, k, conn, newton):
    rp, _, dist = similarity_fixture(nobs, k, conn)
    w, rho, sigma = _gpu_sim(umx, rp, dist, conn=conn, mkds=1e-8, newton=newton)
    ew, erho, esig = oracle.neighbor_similarities(rp, dist, local_connectivity=conn, min_k_dist_scale=1e-8,
                                                  use_newton=newton)
    assert np.array_equal(rho, erho, equal_nan=True)
    _close(sigma, esig, REL)
    _close(w, ew, REL)


def test_similarities_c1(umx):
    _, (rp, _, dist) = c1_knn()
    w, rho, sigma = _gpu_sim(umx, rp, dist)
    ew, erho, esig = oracle.neighbor_similarities(rp, dist)
    assert np.array_equal(rho, erho)
    _close(sigma, esig, REL)
    _close(w, ew, REL)


@pytest.mark.parametrize("conn,bw", [(1.0, 1.0), (1.3, 0.8), (2.5, 1.5)])
def test_similarities_float_input_double_path(umx, conn, bw):
    rp, _, dist = similarity_fixture(200, 20, conn)
    d32 = dist.astype(np.float32)
    w, rho, sigma = _gpu_sim(umx, rp, d32, conn=conn, bw=bw)
    assert w.dtype == np.float32
    # the double path on the same float distances, options stored as float first
    f = lambda v: float(np.float32(v))
    ew, erho, esig = oracle.neighbor_similarities(rp, d32.astype(np.float64), local_connectivity=f(conn),
                                                  bandwidth=f(bw), min_k_dist_scale=f(1e-3))
    assert np.array_equal(rho, erho)
    _close(sigma, esig, REL)
    _close(w, ew.astype(np.float32), REL)


def test_similarities_special_rows(umx):
    # empty, all-zero, all-tied, leading zeros, too-short rows in one launch
    rows = [[], [0.0] * 20, [10.0] * 20, [0.0, 0.0, 0.5, 0.7, 0.9, 1.3], list(np.arange(20) * 0.1),
            [0.3], [0.0, 0.2]]
    rp = np.zeros(len(rows) + 1, np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    dist = np.array([v for r in rows for v in r], dtype=np.float64)
    for conn in (1.0, 100.0, 0.5):
        w, rho, sigma = _gpu_sim(umx, rp, dist, conn=conn)
        ew, erho, esig = oracle.neighbor_similarities(rp, dist, local_connectivity=conn)
        assert np.array_equal(rho, erho, equal_nan=True)
        _close(sigma, esig, REL)
        _close(w, ew, REL)


def test_similarities_bounded_sigma_float(umx):
    rp = np.arange(4, dtype=np.int64) * 20
    d = np.tile(np.arange(20, dtype=np.float32) * np.float32(0.1) + np.float32(0.01), 3).astype(np.float32)
    w, _, _ = _gpu_sim(umx, rp, d, bw=0.0, mkds=0.1, newton=False)
    assert np.all(w <= 1) and np.all(w > 0)
    w0, _, _ = _gpu_sim(umx, rp, d, bw=0.0, mkds=0.0, newton=False)
    for i in range(3):
        row = w0[rp[i]:rp[i + 1]]
        assert row[0] == 1 and np.all(row[1:] == 0)


@pytest.mark.parametrize("k", [3, 33, 64, 65, 130, 300, 513, 700])
def test_similarities_row_lengths(umx, k):
    rng = np.random.default_rng(k)
    n = 97
    dist = np.sort(rng.exponential(1.0, size=(n, k)), axis=1)
    dist[::7, 0] = 0.0
    rp = np.arange(n + 1, dtype=np.int64) * k
    w, rho, sigma = _gpu_sim(umx, rp, dist.reshape(-1), conn=1.3)
    ew, erho, esig = oracle.neighbor_similarities(rp, dist.reshape(-1), local_connectivity=1.3)
    assert np.array_equal(rho, erho)
    _close(sigma, esig, REL)
    _close(w, ew, REL)


def _gpu_combine(umx, rp, idx, val, mix):
    nl = umx.NeighborList(rp, idx, val.copy())
    umx.combine_neighbor_sets(nl, mix)
    return nl.row_ptr, nl.idx, nl.val


@pytest.mark.parametrize("nobs,k", list(itertools.product([50, 100, 200], [5, 10, 15])))
@pytest.mark.parametrize("mix", [1.0, 0.0, 0.5])
def test_combine_reference_grid_bit_exact(umx, nobs, k, mix):
    rp, idx, w = combine_fixture(nobs, k)
    g = _gpu_combine(umx, rp, idx, w, mix)
    e = oracle.combine_neighbor_sets(rp, idx, w, mix)
    for a, b in zip(g, e):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("mix", [1.0, 0.0, 0.3, 0.75])
@pytest.mark.parametrize("seed,n,maxk,hubs", [(1, 300, 12, 3), (2, 1000, 40, 10), (3, 50, 70, 2)])
def test_combine_ragged_hubs_bit_exact(umx, dtype, mix, seed, n, maxk, hubs):
    rp, idx, val = random_ragged_graph(seed, n, maxk, dtype=dtype, hubs=hubs)
    val = np.exp(-val).astype(dtype)
    g = _gpu_combine(umx, rp, idx, val, mix)
    e = oracle.combine_neighbor_sets(rp, idx, val, mix)
    assert g[2].dtype == dtype
    for a, b in zip(g, e):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_combine_mirror_sort_classes_bit_exact(umx, dtype):
    """Rows whose mirrored entries fall in every sorting class of the union:
    up to 32 (32-lane shuffle sort), 33-64 (one wavefront), 65-512 and
    513-8192 (LDS bitonic sorts of two capacities), above 8192 (rank sort in
    global memory).  Hubs with chosen in-degrees are planted in a sparse
    ragged graph."""
    n = 20000
    rp, idx, d = random_ragged_graph(11, n, 6, dtype=dtype, hubs=0)
    rows = [list(zip(idx[rp[i]:rp[i + 1]].tolist(), d[rp[i]:rp[i + 1]].tolist())) for i in range(n)]
    rng = np.random.default_rng(12)
    hubs = {17: 48, 4242: 300, 9001: 2000, 15555: 9000}
    for h, count in hubs.items():
        for i in rng.choice(n, size=count, replace=False):
            i = int(i)
            if i != h and all(t != h for t, _ in rows[i]):
                rows[i].append((h, float(rng.exponential(1.0))))
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    idx = np.array([t for r in rows for t, _ in r], np.int32)
    val = np.exp(-np.array([v for r in rows for _, v in r])).astype(dtype)
    # mirrors of row t: entries i -> t without t -> i
    pairs = set(zip(np.repeat(np.arange(n), np.diff(rp)).tolist(), idx.tolist()))
    mirrors = np.zeros(n, np.int64)
    for i, t in pairs:
        if (t, i) not in pairs:
            mirrors[t] += 1
    for lo, hi in [(1, 32), (33, 64), (65, 512), (513, 8192), (8193, n)]:
        assert np.any((mirrors >= lo) & (mirrors <= hi)), (lo, hi)
    for mix in (1.0, 0.3):
        g = _gpu_combine(umx, rp, idx, val, mix)
        e = oracle.combine_neighbor_sets(rp, idx, val, mix)
        for a, b in zip(g, e):
            assert np.array_equal(a, b)


def test_combine_c1_after_similarities(umx):
    _, (rp, idx, dist) = c1_knn()
    ew, _, _ = oracle.neighbor_similarities(rp, dist)
    g = _gpu_combine(umx, rp, idx, ew, 1.0)
    e = oracle.combine_neighbor_sets(rp, idx, ew, 1.0)
    for a, b in zip(g, e):
        assert np.array_equal(a, b)


def test_combine_rejects_bad_graphs(umx):
    rp = np.array([0, 2, 3], dtype=np.int64)
    for idx in ([1, 1, 0], [0, 1, 0], [1, 2, 0], [1, -1, 0]):
        with pytest.raises(umx.UmxError) as ei:
            _gpu_combine(umx, rp, np.array(idx, np.int32), np.ones(3), 1.0)
        assert ei.value.code == -2


@pytest.mark.parametrize("nobs,k", [(50, 5), (100, 10), (200, 15)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("num_epochs,nsr", [(500, 5.0), (200, 3.0), (7, 0.5)])
def test_epochs_bit_exact(umx, nobs, k, dtype, num_epochs, nsr):
    _, rp, idx, w = layout_fixture(nobs, k)
    w = w.astype(dtype)
    w[0] = dtype(1e-8)
    nl = umx.NeighborList(rp, idx, w)
    ed = umx.similarities_to_epochs(nl, num_epochs, nsr)
    e = oracle.similarities_to_epochs(rp, idx, w, num_epochs, nsr)
    assert np.array_equal(ed.cumulative_num_edges, e["row_ptr"])
    assert np.array_equal(ed.edge_targets, e["targets"])
    assert np.array_equal(ed.epochs_per_sample, e["eps"])
    assert np.array_equal(ed.epoch_of_next_sample, e["next"])
    assert np.array_equal(ed.epoch_of_next_negative_sample, e["next_neg"])
    assert ed.total_epochs == num_epochs and ed.current_epoch == 0
    assert np.all(ed.epochs_per_sample >= 1)


def test_full_graph_pipeline_resident(umx):
    """Similarities -> union -> schedule on one resident graph equals the
    oracle's pipeline (C1)."""
    _, (rp, idx, dist) = c1_knn()
    g = umx.Graph(umx.NeighborList(rp, idx, dist))
    g.neighbor_similarities()
    ew, _, _ = oracle.neighbor_similarities(rp, dist)
    gw = g.export()
    # union inputs differ only by the similarity rounding; compare the union of
    # the GPU weights on both sides for bit-exactness
    _close(umx.Graph(umx.NeighborList(rp, idx, dist)).export().val, dist, 0)
    g.combine(1.0)
    e = oracle.combine_neighbor_sets(rp, idx, gw.val, 1.0)
    got = g.export()
    assert np.array_equal(got.row_ptr, e[0]) and np.array_equal(got.idx, e[1]) and np.array_equal(got.val, e[2])
    lay = umx.Layout(graph=g, num_epochs=500, negative_sample_rate=5.0)
    ed = lay.epoch_data()
    eo = oracle.similarities_to_epochs(e[0], e[1], e[2], 500, 5.0)
    assert np.array_equal(ed.epochs_per_sample, eo["eps"]) and np.array_equal(ed.edge_targets, eo["targets"])
