"""GPU spectral initialisation and connected components against the oracle.

* mock graphs of the reference's spectral test (spectral_init.cpp:87-128,
  N in {50, 100, 200} x num_dim in {2, 5}) and config C1's graph: the GPU
  coordinates equal the dense oracle's (the same sign convention) within
  1e-6 of the scale, its eigenvalues those of a dense eigen-solve, and the
  reference's own checks hold (eigenvectors of L, max |value| = scale, no
  zeros);
* several components: failure with the embedding untouched
  (spectral_init.cpp:130-149), and the component count equals scipy's;
* jitter adds noise of the requested size;
* initialize() with SPECTRAL uses it, and falls back to RANDOM on failure;
* a 30k-point connected kNN graph, where the Lanczos run is extended
  several times before its Ritz pairs converge: eigenvalues equal scipy's
  ARPACK (a different eigensolver) within 1e-8, and every output column is
  an eigenvector of 2I - L (relative residual at float precision).
"""
import numpy as np
import pytest
import scipy.sparse
import scipy.sparse.csgraph

import oracle
from _data import c1_knn
from test_oracle_spectral import check_eigenvectors, mock_probabilities

pytestmark = pytest.mark.gpu


def _graph(umx, rp, idx, val):
    """A resident graph holding an already symmetric weighted CSR (the union
    step leaves a symmetric graph unchanged at mix 1 only if weights are
    symmetric, so this loads it directly)."""
    return umx.Graph(umx.NeighborList(rp, idx, val))


@pytest.mark.parametrize("n", [50, 100, 200])
@pytest.mark.parametrize("d", [2, 5])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_mock_graphs_match_dense_oracle(umx, n, d, dtype):
    rp, idx, val = mock_probabilities(n, n)
    g = _graph(umx, rp, idx, val.astype(dtype))
    assert g.components() == 1
    ok, Y, ev = g.spectral_init(d, scale=10.0)
    assert ok
    okr, Yr, wr = oracle.spectral_init_dense(rp, idx, val.astype(dtype).astype(np.float64), d, scale=10.0)
    assert okr
    tol = 1e-6 if dtype == np.float64 else 2e-5
    assert np.max(np.abs(Y.astype(np.float64) - Yr)) <= tol * 10.0
    assert np.allclose(ev, wr, atol=1e-9)
    assert np.all(Y != 0)
    assert np.isclose(np.max(np.abs(Y)), 10.0, rtol=1e-6)
    check_eigenvectors(rp, idx, val.astype(dtype).astype(np.float64), Y.astype(np.float64))


def test_c1_graph_matches_dense_oracle(umx):
    _, (rp, idx, dist) = c1_knn()
    g = umx.Graph(umx.NeighborList(rp, idx, dist.copy()))
    g.neighbor_similarities()
    g.combine(1.0)
    nl = g.export()
    ok, Y, ev = g.spectral_init(2, scale=10.0)
    assert ok
    okr, Yr, wr = oracle.spectral_init_dense(nl.row_ptr, nl.idx, nl.val, 2, scale=10.0)
    assert okr
    print(f"C1 spectral: eigenvalues of 2I - L {ev}, max |GPU - dense| {np.max(np.abs(Y - Yr)):.2e}")
    assert np.max(np.abs(Y - Yr)) <= 1e-6 * 10.0
    assert np.allclose(ev, wr, atol=1e-9)


def test_multiple_components_fail_and_leave_embedding(umx):
    rp1, i1, v1 = mock_probabilities(50, 1)
    rp2, i2, v2 = mock_probabilities(100, 2)
    rp = np.concatenate([rp1, rp2[1:] + rp1[-1]])
    idx = np.concatenate([i1, i2 + 50]).astype(np.int32)
    val = np.concatenate([v1, v2])
    g = _graph(umx, rp, idx, val)
    assert g.components() == 2
    out = np.full((150, 2), 7.0)
    ok, Y, _ = g.spectral_init(2, out=out)
    assert not ok and Y is None
    assert np.all(out == 7.0)


@pytest.mark.parametrize("seed", range(4))
def test_component_count_vs_scipy(umx, seed):
    n = 2000
    A = scipy.sparse.random(n, n, density=0.0012, random_state=seed, format="csr")
    A = (A + A.T).tocsr()
    A.setdiag(0)
    A.eliminate_zeros()
    A.sort_indices()
    want, _ = scipy.sparse.csgraph.connected_components(A, directed=False)
    g = _graph(umx, A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data)
    assert g.components() == want


def test_jitter(umx):
    rp, idx, val = mock_probabilities(200, 3)
    g = _graph(umx, rp, idx, val)
    ok, Y0, _ = g.spectral_init(2)
    ok1, Y1, _ = g.spectral_init(2, jitter=True, jitter_sd=0.01, seed=5)
    assert ok and ok1
    diff = (Y1 - Y0).reshape(-1)
    assert abs(diff.mean()) < 0.003 and 0.008 < diff.std() < 0.012
    ok2, Y2, _ = g.spectral_init(2, jitter=True, jitter_sd=0.01, seed=5)
    assert np.array_equal(Y1, Y2)


def test_initialize_spectral_and_fallback(umx):
    _, (rp, idx, dist) = c1_knn()
    nl = umx.NeighborList(rp, idx, dist.copy())
    emb = np.zeros((2000, 2))
    opt = umx.Options()
    st = umx.initialize(nl, 2, emb, opt)
    assert np.isclose(np.max(np.abs(emb)), 10.0)
    g = umx.Graph(umx.NeighborList(rp, idx, dist.copy()))
    g.neighbor_similarities()
    g.combine(1.0)
    _, Y, _ = g.spectral_init(2)
    assert np.allclose(emb, Y)
    st.run(emb, 5)
    assert np.all(np.isfinite(emb))
    # two components: falls back to random coordinates in (-10, 10)
    rp2 = np.concatenate([rp, rp[1:] + rp[-1]])
    idx2 = np.concatenate([idx, idx + 2000]).astype(np.int32)
    d2 = np.concatenate([dist, dist])
    emb2 = np.zeros((4000, 2))
    umx.initialize(umx.NeighborList(rp2, idx2, d2), 2, emb2, umx.Options())
    assert np.all(np.abs(emb2) <= 10.0) and np.std(emb2) > 4.0


def test_extended_lanczos_matches_arpack(umx):
    import scipy.sparse.linalg

    n, k = 30_000, 15
    x = np.random.default_rng(31).normal(size=(n, 20)).astype(np.float32)
    idx, dist = umx.find_knn(x, k)
    g = umx.Graph(umx.NeighborList(np.arange(n + 1, dtype=np.int64) * k, idx.reshape(-1).copy(),
                                   dist.reshape(-1).copy()))
    g.neighbor_similarities()
    g.combine(1.0)
    ok, Y, ev = g.spectral_init(2, scale=10.0)
    rep = g.spectral_report()
    print(f"30k spectral: eigenvalues {ev}, {rep}")
    assert ok and rep["attempts"] >= 2 and rep["worst_residual"] <= 1e-6 and rep["bytes"] > 0
    nl = g.export()
    W = scipy.sparse.csr_matrix((nl.val.astype(np.float64), nl.idx, nl.row_ptr), shape=(n, n))
    isd = 1.0 / np.sqrt(np.asarray(W.sum(axis=1)).ravel())
    M = scipy.sparse.identity(n, format="csr") + scipy.sparse.diags(isd) @ W @ scipy.sparse.diags(isd)
    w = np.sort(scipy.sparse.linalg.eigsh(M, k=3, which="LA", tol=1e-10)[0])[::-1]
    assert np.max(np.abs(w - ev) / w) <= 1e-8
    assert np.isclose(np.max(np.abs(Y)), 10.0, rtol=1e-6)
    for c in range(2):
        y = Y[:, c].astype(np.float64)
        assert np.linalg.norm(M @ y - ev[c + 1] * y) / np.linalg.norm(y) <= 1e-5
