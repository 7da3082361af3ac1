"""Pins the oracle's neighbor_similarities (CPU only).

Checks follow the reference's own tests (tests/src/neighbor_similarities.cpp)
plus an independent root check: for each row, the target equation
sum_k exp(-max(0, d_k - rho) / sigma) = log2(k + 1) * bandwidth is re-evaluated
in numpy at the oracle's (rho, sigma), and rho is recomputed from its
definition (interpolated distance to the local_connectivity-th non-zero
neighbour).  A wrong rho index, a dropped count of d <= rho or a wrong target
fails one of them.
"""
import itertools
import math

import numpy as np
import pytest

import oracle
from _data import similarity_fixture

GRID = list(itertools.product([50, 100, 200], [5, 10, 20], [0.4, 1.0, 1.3, 2.0, 2.5]))


def _check_row_invariants(rp, w, conn, tol=1e-5):
    for i in range(rp.size - 1):
        s = w[rp[i]:rp[i + 1]]
        k = s.size
        for q in range(k):
            if q < int(conn):
                assert s[q] == 1.0
            else:
                assert s[q] < 1.0
        assert np.all(np.diff(s) <= 0)
        assert abs(s.sum() - math.log2(k + 1)) < tol


@pytest.mark.parametrize("nobs,k,conn", GRID)
def test_reference_grid_newton_and_bisection(nobs, k, conn):
    rp, _, dist = similarity_fixture(nobs, k, conn)
    w, rho, sigma = oracle.neighbor_similarities(rp, dist, local_connectivity=conn, min_k_dist_scale=1e-8)
    _check_row_invariants(rp, w, conn)
    # thread-count invariance (neighbor_similarities.cpp:82-87)
    w3, _, _ = oracle.neighbor_similarities(rp, dist, local_connectivity=conn, min_k_dist_scale=1e-8, nthreads=3)
    assert np.array_equal(w, w3)
    # bisection-only mode (neighbor_similarities.cpp:90-118)
    wb, _, _ = oracle.neighbor_similarities(rp, dist, local_connectivity=conn, min_k_dist_scale=1e-8,
                                            use_newton=False)
    _check_row_invariants(rp, wb, conn)


@pytest.mark.parametrize("nobs,k,conn", [(100, 10, 0.4), (200, 20, 1.3), (50, 5, 2.5), (200, 10, 1.0)])
def test_rho_definition_and_root(nobs, k, conn):
    rp, _, dist = similarity_fixture(nobs, k, conn)
    w, rho, sigma = oracle.neighbor_similarities(rp, dist, local_connectivity=conn, min_k_dist_scale=1e-8)
    raw = int(conn)
    frac = conn - raw
    for i in range(nobs):
        d = dist[rp[i]:rp[i + 1]]
        nz = int(np.argmax(d != 0)) if np.any(d != 0) else d.size
        ci = nz + raw
        lower = d[ci - 1] if ci > 0 else 0.0
        expect_rho = lower + frac * (d[ci] - lower)
        assert rho[i] == pytest.approx(expect_rho, rel=0, abs=1e-15)
        obs = np.sum(np.exp(-np.maximum(0.0, d - rho[i]) / sigma[i]))
        assert abs(obs - math.log2(k + 1)) < 1e-5
        # weights are the kernel at (rho, sigma)
        expect_w = np.where(d > rho[i], np.exp(-(d - rho[i]) / sigma[i]), 1.0)
        np.testing.assert_allclose(w[rp[i]:rp[i + 1]], expect_w, rtol=1e-13)


def test_empty_row():
    rp = np.array([0, 0], dtype=np.int64)
    w, rho, sigma = oracle.neighbor_similarities(rp, np.zeros(0))
    assert w.size == 0 and np.isnan(rho[0]) and np.isnan(sigma[0])


def test_all_zero_distance():
    rp = np.arange(4, dtype=np.int64) * 20
    w, _, _ = oracle.neighbor_similarities(rp, np.zeros(60))
    assert np.all(w == 1)


def test_no_distance_above_rho():
    rp = np.arange(4, dtype=np.int64) * 20
    w, _, _ = oracle.neighbor_similarities(rp, np.full(60, 10.0))
    assert np.all(w == 1)


def test_too_high_connectivity():
    rp = np.arange(4, dtype=np.int64) * 20
    d = np.tile(np.arange(20) * 0.1, 3)
    w, _, _ = oracle.neighbor_similarities(rp, d, local_connectivity=100.0)
    assert np.all(w == 1)


def test_bounded_sigma_float_bisection():
    rp = np.arange(4, dtype=np.int64) * 20
    d = np.tile(np.arange(20, dtype=np.float32) * np.float32(0.1) + np.float32(0.01), 3).astype(np.float32)
    w, _, _ = oracle.neighbor_similarities(rp, d, bandwidth=0.0, min_k_dist_scale=0.1, use_newton=False)
    assert w.dtype == np.float32
    assert np.all(w <= 1) and np.all(w > 0)
    w0, _, _ = oracle.neighbor_similarities(rp, d, bandwidth=0.0, min_k_dist_scale=0.0, use_newton=False)
    for i in range(3):
        row = w0[rp[i]:rp[i + 1]]
        assert row[0] == 1 and np.all(row[1:] == 0)


def test_float_path_is_float_arithmetic():
    # Float_ = float computes in float: its result differs from the double path
    # only at float rounding level.
    rp, _, dist = similarity_fixture(100, 10, 1.0)
    wd, _, _ = oracle.neighbor_similarities(rp, dist)
    wf, _, _ = oracle.neighbor_similarities(rp, dist.astype(np.float32))
    np.testing.assert_allclose(wf, wd, rtol=2e-4, atol=1e-6)
