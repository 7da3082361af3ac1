"""Config C2 at full size: 100,000 observations x 50 dims f32 from a
30-component Gaussian mixture, exact kNN k = 15, out_dim 2, 200 epochs.

One graph for both paths: the kNN lists come from the library's exact GPU
brute force (bit-exact against the oracle's brute force at smaller sizes,
test_gpu_knn.py).  From them:

* the GPU's fuzzy union (f32 input, solve in double) has the oracle's
  structure bit for bit (row lengths, indices, order) and its weights within
  1e-6 relative of the oracle's double path (the reference's Float_ =
  double path; its float path solves in float, about 4e-5 away);
* the GPU layout (production pull kernel, counter sampler) and the oracle's
  serial layout (the reference's loop and mt19937_64 sampler), from the same
  initial embedding and seed, agree on 15-NN preservation and
  trustworthiness within 0.01 (both metrics on 1,500 seeded query points).
"""
import numpy as np
import pytest

import oracle
from _data import mixture
from _quality import knn_preservation, trustworthiness

pytestmark = pytest.mark.gpu


def test_c2_graph_and_layout_quality(umx):
    n, k, T, nsr, seed = 100_000, 15, 200, 5.0, 1234567890
    x, _ = mixture(2, n, 50, 30)
    idx, dist = umx.find_knn(x, k)
    rp = np.arange(n + 1, dtype=np.int64) * k
    # fuzzy graph: oracle (reference algorithm, double path) and GPU
    w_o, _, _ = oracle.neighbor_similarities(rp, dist.reshape(-1).astype(np.float64), nthreads=8)
    crp, cidx, cw = oracle.combine_neighbor_sets(rp, idx.reshape(-1), w_o, 1.0)
    g = umx.Graph(umx.NeighborList(rp, idx.reshape(-1).copy(), dist.reshape(-1).copy()))
    g.neighbor_similarities()
    g.combine(1.0)
    nl = g.export()
    assert np.array_equal(nl.row_ptr, crp) and np.array_equal(nl.idx, cidx)
    assert np.max(np.abs(nl.val.astype(np.float64) - cw) / np.maximum(np.abs(cw), 1e-30)) <= 1e-6

    a, b = umx.find_ab(1.0, 0.1)
    emb0 = np.random.default_rng(2024).uniform(-10, 10, size=(n, 2)).astype(np.float32)
    eo = emb0.copy()
    ep = oracle.similarities_to_epochs(crp, cidx, cw.astype(np.float32), T, nsr)
    oracle.Layout(ep, T, nsr, rng_kind=oracle.RNG_MT, seed=seed, dtype=np.float32).run(eo, a, b)
    lay = umx.Layout(graph=g, num_epochs=T, negative_sample_rate=nsr)
    lay.set_params(2, a, b, 1.0, 1.0, seed)
    lay.set_embedding(emb0)
    lay.run(T)
    eg = lay.get_embedding()

    q = np.random.default_rng(5).choice(n, 1500, replace=False)
    hi = idx.reshape(n, k)
    p_cpu = knn_preservation(x, eo, 15, high_idx=hi, queries=q)
    p_gpu = knn_preservation(x, eg, 15, high_idx=hi, queries=q)
    t_cpu = trustworthiness(x, eo, 15, queries=q)
    t_gpu = trustworthiness(x, eg, 15, queries=q)
    print(f"C2 100k: 15-NN preservation oracle {p_cpu:.4f} gpu {p_gpu:.4f}; "
          f"trustworthiness oracle {t_cpu:.4f} gpu {t_gpu:.4f}")
    assert abs(p_cpu - p_gpu) <= 0.01 and abs(t_cpu - t_gpu) <= 0.01
