"""The quality metrics used for final-embedding parity, checked against
sklearn on small inputs (CPU only)."""
import numpy as np
import pytest

from _quality import knn_preservation, trustworthiness


def test_trustworthiness_matches_sklearn():
    sk = pytest.importorskip("sklearn.manifold")
    rng = np.random.default_rng(0)
    x = rng.normal(size=(300, 8))
    y = x[:, :2] + 0.3 * rng.normal(size=(300, 2))
    for k in (5, 15):
        assert trustworthiness(x, y, k) == pytest.approx(sk.trustworthiness(x, y, n_neighbors=k), abs=1e-12)


def test_preservation_extremes():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(200, 5))
    assert knn_preservation(x, x, 10) == 1.0
    y = rng.normal(size=(200, 2))
    assert knn_preservation(x, y, 10) < 0.2
