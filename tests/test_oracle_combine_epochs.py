"""Pins the oracle's combine_neighbor_sets and similarities_to_epochs (CPU only).

combine: the reference's brute-force symmetrisation (combine_neighbor_sets.cpp:
47-112) restated in Python over a dict of unordered pairs — a different
algorithm from the cursor walk — plus symmetry, sortedness and the edge-count
inequalities of the reference test (:128-148).
epochs: the closed form eps = max / w over the survivors w >= max / num_epochs
and the reference's Epochs test (optimize_layout.cpp:58-77).
"""
import itertools

import numpy as np
import pytest

import oracle
from _data import combine_fixture, layout_fixture, random_ragged_graph


def slow_symmetrization(rp, idx, val, mix):
    probs = {}
    for i in range(rp.size - 1):
        for e in range(rp[i], rp[i + 1]):
            j, z = int(idx[e]), float(val[e])
            key = (max(i, j), min(i, j))
            if key in probs:
                prev = -probs[key]
                if mix == 1:
                    cur = prev + z - z * prev
                elif mix == 0:
                    cur = prev * z
                else:
                    prod = prev * z
                    cur = mix * (z + prev - prod) + (1 - mix) * prod
                probs[key] = cur
            else:
                probs[key] = -z
    out = {}
    for key, v in probs.items():
        if v < 0 or (v == 0 and np.signbit(v)):
            if mix == 0:
                continue
            v = -v if mix == 1 else -v * mix
        out[key] = v
    return out


def compare_to_slow(rp, idx, val, orp, oidx, oval, mix, rtol):
    expect = slow_symmetrization(rp, idx, val, mix)
    found = set()
    for i in range(orp.size - 1):
        for e in range(orp[i], orp[i + 1]):
            j = int(oidx[e])
            key = (max(i, j), min(i, j))
            assert key in expect
            assert oval[e] == pytest.approx(expect[key], rel=rtol, abs=0)
            found.add(key)
    assert len(found) == len(expect)


def check_structure(orp, oidx, oval):
    n = orp.size - 1
    lookup = {}
    for i in range(n):
        row = oidx[orp[i]:orp[i + 1]]
        assert np.all(np.diff(row) > 0), "rows sorted by index, unique"
        assert not np.any(row == i)
        for e in range(orp[i], orp[i + 1]):
            lookup[(i, int(oidx[e]))] = oval[e]
    for (i, j), v in lookup.items():
        assert (j, i) in lookup and lookup[(j, i)] == v, "symmetric weights, bit for bit"


@pytest.mark.parametrize("nobs,k", list(itertools.product([50, 100, 200], [5, 10, 15])))
def test_reference_grid(nobs, k):
    rp, idx, w = combine_fixture(nobs, k)
    res = {}
    for mix in (1.0, 0.0, 0.5):
        orp, oidx, oval = oracle.combine_neighbor_sets(rp, idx, w, mix)
        compare_to_slow(rp, idx, w, orp, oidx, oval, mix, rtol=1e-12)
        check_structure(orp, oidx, oval)
        res[mix] = orp
    du, di, dm, do = (np.diff(res[1.0]), np.diff(res[0.0]), np.diff(res[0.5]), np.diff(rp))
    assert np.all(du == dm) and np.all(du >= di) and np.all(du >= do) and np.all(di <= do)
    assert du.sum() > di.sum() and du.sum() > do.sum() and di.sum() < do.sum()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("mix", [1.0, 0.0, 0.3])
def test_ragged_with_hubs(dtype, mix):
    rp, idx, val = random_ragged_graph(11, 300, 12, dtype=dtype, hubs=3)
    val = np.exp(-val).astype(dtype)
    orp, oidx, oval = oracle.combine_neighbor_sets(rp, idx, val, mix)
    assert oval.dtype == dtype
    compare_to_slow(rp, idx, val.astype(np.float64), orp, oidx, oval.astype(np.float64), mix,
                    rtol=1e-12 if dtype == np.float64 else 3e-7)
    check_structure(orp, oidx, oval)


def test_rejects_self_and_out_of_range():
    rp = np.array([0, 1, 2], dtype=np.int64)
    with pytest.raises(ValueError):
        oracle.combine_neighbor_sets(rp, np.array([0, 0], np.int32), np.ones(2), 1.0)
    with pytest.raises(ValueError):
        oracle.combine_neighbor_sets(rp, np.array([1, 5], np.int32), np.ones(2), 1.0)


@pytest.mark.parametrize("nobs,k", [(50, 5), (100, 10), (200, 15)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_epochs_closed_form(nobs, k, dtype):
    _, rp, idx, w = layout_fixture(nobs, k)
    w = w.astype(dtype)
    w[0] = dtype(1e-8)  # forced removal (optimize_layout.cpp:59)
    ep = oracle.similarities_to_epochs(rp, idx, w, 500, 5.0)
    maxed = w.max()
    assert ep["maxed"] == maxed
    limit = dtype(maxed / dtype(500))
    keep = w >= limit
    assert not keep[0]
    assert ep["targets"].size == int(keep.sum()) < w.size
    np.testing.assert_array_equal(ep["targets"], idx[keep])
    counts = np.add.reduceat(keep.astype(np.int64), rp[:-1]) if w.size else np.zeros(nobs)
    counts[np.diff(rp) == 0] = 0
    np.testing.assert_array_equal(np.diff(ep["row_ptr"]), counts)
    expect_eps = (maxed / w[keep]).astype(dtype)
    np.testing.assert_array_equal(ep["eps"], expect_eps)
    np.testing.assert_array_equal(ep["next"], expect_eps)
    np.testing.assert_array_equal(ep["next_neg"], (expect_eps / dtype(5.0)).astype(dtype))
    assert np.all(ep["eps"] >= 1)
