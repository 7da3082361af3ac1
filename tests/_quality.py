"""Embedding quality metrics used to compare final layouts.

knn_preservation: mean fraction of each point's k nearest neighbours in the
input space that are also among its k nearest neighbours in the embedding.
trustworthiness: Venna & Kaski's measure, the formula sklearn implements
(sklearn.manifold.trustworthiness), cross-checked on small inputs in
tests/test_quality_metrics.py; computed here on an optional subsample of
query points with exact ranks over all points.
"""
import numpy as np


def _knn_numpy(x, k, queries=None, block=512):
    x = np.asarray(x, dtype=np.float64)
    q = x if queries is None else x[queries]
    qi = np.arange(x.shape[0]) if queries is None else np.asarray(queries)
    sq = np.sum(x * x, axis=1)
    out = np.empty((q.shape[0], k), dtype=np.int64)
    for s in range(0, q.shape[0], block):
        qq = q[s:s + block]
        d = np.sum(qq * qq, axis=1)[:, None] + sq[None, :] - 2.0 * qq @ x.T
        d[np.arange(qq.shape[0]), qi[s:s + block]] = np.inf
        part = np.argpartition(d, k, axis=1)[:, :k]
        dd = np.take_along_axis(d, part, axis=1)
        order = np.argsort(dd, axis=1, kind="stable")
        out[s:s + block] = np.take_along_axis(part, order, axis=1)
    return out


def knn_preservation(x_high, y_low, k=15, high_idx=None, queries=None):
    hi = high_idx if high_idx is not None else _knn_numpy(x_high, k, queries)
    if queries is not None and high_idx is not None:
        hi = hi[queries]
    lo = _knn_numpy(y_low, k, queries)
    hits = [len(set(a[:k]).intersection(b[:k])) for a, b in zip(hi, lo)]
    return float(np.mean(hits)) / k


def trustworthiness(x_high, y_low, k=15, queries=None):
    x = np.asarray(x_high, dtype=np.float64)
    n = x.shape[0]
    qi = np.arange(n) if queries is None else np.asarray(queries)
    lo = _knn_numpy(y_low, k, qi)
    sq = np.sum(x * x, axis=1)
    penalty = 0.0
    for s in range(0, qi.size, 256):
        ids = qi[s:s + 256]
        d = sq[ids][:, None] + sq[None, :] - 2.0 * x[ids] @ x.T
        d[np.arange(ids.size), ids] = -np.inf  # self ranks 0
        order = np.argsort(d, axis=1, kind="stable")
        ranks = np.empty_like(order)
        ranks[np.arange(ids.size)[:, None], order] = np.arange(n)[None, :]
        r = np.take_along_axis(ranks, lo[s:s + 256], axis=1)  # 1-based ranks among non-self
        penalty += np.sum(np.maximum(r - k, 0))
    m = qi.size
    return 1.0 - 2.0 / (m * k * (2.0 * n - 3.0 * k - 1.0)) * penalty
