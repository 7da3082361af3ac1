"""Pins the spectral-initialisation oracle and the host eigen-solver.

* the reference's eigenvector check (spectral_init.cpp:36-85): the output
  columns are eigenvectors of the normalized Laplacian, max |value| = scale,
  no zero entries; their eigenvalues are the smallest positive ones, found
  independently from L's own spectrum;
* the reference's component known answers (spectral_init.cpp:160-200) and
  scipy's connected_components on random graphs;
* a multi-component graph fails (spectral_init.cpp:130-149);
* the library's tridiagonal QR solver (host code) against
  scipy.linalg.eigh_tridiagonal.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg
import scipy.sparse
import scipy.sparse.csgraph

import oracle


def mock_probabilities(n, seed):
    """Structure of spectral_init.cpp:14-34 (a chain forcing one component
    plus each pair with probability 0.2, symmetric uniform weights); numpy
    draws stand in for the reference's aarand uniforms."""
    rng = np.random.default_rng(seed)
    rows = [[] for _ in range(n)]
    for r in range(n):
        forced = (n if r == 0 else r) - 1
        for c in range(r):
            if c == forced or rng.random() < 0.2:
                v = rng.random()
                rows[r].append((c, v))
                rows[c].append((r, v))
    for r in rows:
        r.sort()
    rp = np.zeros(n + 1, dtype=np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    idx = np.array([c for r in rows for c, _ in r], dtype=np.int32)
    val = np.array([v for r in rows for _, v in r])
    return rp, idx, val


def check_eigenvectors(rp, idx, val, Y):
    """spectral_init.cpp:36-85 restated: L Y is a per-column rescaling of Y."""
    n, d = Y.shape
    sums = np.sqrt(np.add.reduceat(val, rp[:-1]) if val.size else np.zeros(n))
    copy = Y.copy()
    rows = np.repeat(np.arange(n), np.diff(rp))
    np.add.at(copy, rows, (-val / sums[rows] / sums[idx])[:, None] * Y[idx])
    res_scale = Y.sum(axis=0)
    copy_scale = copy.sum(axis=0)
    scaled = copy * res_scale / copy_scale
    assert np.all(np.abs(Y - scaled) < 1e-4 * (np.abs(Y) + np.abs(scaled)) / 2)


@pytest.mark.parametrize("n", [50, 100, 200])
@pytest.mark.parametrize("d", [2, 5])
def test_dense_oracle_reference_invariants(n, d):
    rp, idx, val = mock_probabilities(n, n)
    ok, Y, w = oracle.spectral_init_dense(rp, idx, val, d, scale=10.0)
    assert ok
    assert np.all(Y != 0)
    assert np.isclose(np.max(np.abs(Y)), 10.0, rtol=1e-12)
    check_eigenvectors(rp, idx, val, Y)
    # eigenvalues: the 2nd .. (d+1)-th smallest of L, from L's own spectrum
    W = scipy.sparse.csr_matrix((val, idx, rp), shape=(n, n)).toarray()
    isd = 1 / np.sqrt(W.sum(1))
    L = np.eye(n) - isd[:, None] * W * isd[None, :]
    lam = np.sort(np.linalg.eigvalsh(L))
    assert np.allclose(2.0 - w, lam[:d + 1], atol=1e-10)
    Ln = L @ (Y / np.linalg.norm(Y, axis=0))
    got = np.sum(Ln * (Y / np.linalg.norm(Y, axis=0)), axis=0)
    assert np.allclose(got, lam[1:d + 1], atol=1e-10)


def _lists_to_csr(rows):
    rp = np.zeros(len(rows) + 1, dtype=np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    idx = np.array([c for r in rows for c in r], dtype=np.int32)
    return rp, idx


def _symmetrize(rows):
    out = [list(r) for r in rows]
    for i, r in enumerate(rows):
        for c in r:
            out[c].append(i)
    return out


def test_component_known_answers():
    """spectral_init.cpp:160-200."""
    edges = [[], [], [], [2], [0, 1]]
    assert oracle.has_multiple_components(*_lists_to_csr(_symmetrize(edges)))
    edges[3].append(1)
    assert not oracle.has_multiple_components(*_lists_to_csr(_symmetrize(edges)))
    empty = [[] for _ in range(5)]
    assert oracle.has_multiple_components(*_lists_to_csr(empty))
    chain = [[], [0], [1], [2], [3]]
    assert not oracle.has_multiple_components(*_lists_to_csr(_symmetrize(chain)))


@pytest.mark.parametrize("seed", range(5))
def test_components_vs_scipy(seed):
    rng = np.random.default_rng(seed)
    n = 300
    A = scipy.sparse.random(n, n, density=0.004, random_state=seed, format="csr")
    A = ((A + A.T) > 0).astype(np.float64).tocsr()
    A.setdiag(0)
    A.eliminate_zeros()
    ncomp, _ = scipy.sparse.csgraph.connected_components(A, directed=False)
    assert oracle.has_multiple_components(A.indptr, A.indices) == (ncomp != 1)
    del rng


def test_multiple_components_fail():
    rp1, i1, v1 = mock_probabilities(50, 1)
    rp2, i2, v2 = mock_probabilities(100, 2)
    rp = np.concatenate([rp1, rp2[1:] + rp1[-1]])
    idx = np.concatenate([i1, i2 + 50])
    val = np.concatenate([v1, v2])
    ok, Y, _ = oracle.spectral_init_dense(rp, idx, val, 2)
    assert not ok and Y is None


@pytest.mark.parametrize("k", [1, 2, 3, 8, 60, 300, 512])
def test_tridiagonal_solver_vs_scipy(umx, k):
    rng = np.random.default_rng(k)
    for trial in range(5):
        d = rng.normal(size=k)
        e = rng.normal(size=max(k - 1, 0))
        if trial == 2 and k > 3:
            d[:] = 1.0  # clustered spectrum
            e *= 1e-3
        if trial == 3 and k > 3:
            e[k // 2] = 0.0  # already split into two blocks
            e[0] = 0.0
        if trial == 4:
            d = 2.0 - np.abs(d) * 1e-3 ** np.linspace(0, 1, k)  # graded, near 2 like the Lanczos T
        dd = d.copy()
        z = np.zeros((k, k))
        rc = umx.lib().umx_tridiagonal_eigen(k, dd.ctypes.data_as(C.c_void_p),
                                              e.ctypes.data_as(C.c_void_p) if k > 1 else None,
                                              z.ctypes.data_as(C.c_void_p))
        assert rc == 0
        ref = scipy.linalg.eigh_tridiagonal(d, e, eigvals_only=True)
        assert np.allclose(np.sort(dd), ref, atol=1e-12 * max(1.0, np.max(np.abs(ref))))
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        assert np.allclose(z.T @ z, np.eye(k), atol=1e-10)
        assert np.allclose(z @ np.diag(dd) @ z.T, T, atol=1e-10)
