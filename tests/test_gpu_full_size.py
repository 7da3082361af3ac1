"""The full-size workloads on the GPU, through the C ABI.

1M points (the headline workload: 1,000,000 x 50 float32 from a
100-component Gaussian mixture, exact kNN k = 15, D = 2, 200 epochs):

* rho equal to the oracle's double path and sigma and the fuzzy weights
  within 1e-6 relative of it;
* the fuzzy union's structure (row lengths, indices, order) bit for bit,
  its weights within 1e-6 relative;
* the epoch schedule of the GPU's graph bit for bit, and after a captured
  block of 10 epochs of the production kernel (more than 4096 chunks, so its
  automatic 16-column tile) epoch_of_next_sample,
  epoch_of_next_negative_sample and the per-epoch tallies bit for bit against
  the oracle's counter-sampler run;
* every step among the first 2,000,000 trace records of epoch 10 equal to
  the reference's update from its traced inputs within 1e-5 relative, head
  and negative-sample steps reading other points of earlier phases as the
  epoch left them and all others as they were at its start.

64 instances (sizes log-uniform in [2,000, 200,000], seed 4; each 30 dims
float32 from its own 10-component mixture, seed 4000 + i; kNN k = 15;
200 epochs each):

* every instance's fuzzy graph from the batched build equal to the oracle's
  (structure bit for bit, weights within 1e-6), and its schedule after 3
  epochs of the batched run bit for bit against the oracle's run of that
  instance alone;
* the batched run at least 3 times faster than 64 sequential plain runs.
"""
import time

import numpy as np
import pytest

import oracle
from _data import mixture
from _steps import phase_of

pytestmark = pytest.mark.gpu

NSR = 5.0
SEED = 1234567890


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))) if a.size else 0.0


def test_headline_1m_graph_schedule_and_steps(umx):
    n, k, T = 1_000_000, 15, 200
    x, _ = mixture(3, n, 50, 100)
    idx, dist = umx.find_knn(x, k)
    del x
    rp = np.arange(n + 1, dtype=np.int64) * k
    idx = idx.reshape(-1)
    dist = dist.reshape(-1)
    print(f"1M: kNN done, {idx.size} directed entries", flush=True)

    # K1: similarities, oracle double path vs GPU (solve in double)
    w_o, rho_o, sig_o = oracle.neighbor_similarities(rp, dist.astype(np.float64), nthreads=8)
    g = umx.Graph(umx.NeighborList(rp, idx.copy(), dist.copy()))
    g.neighbor_similarities()
    rho_g, sig_g = g.rho_sigma()
    assert np.array_equal(rho_g, rho_o)
    assert _rel(sig_g, sig_o) <= 1e-6
    assert _rel(g.export().val, w_o) <= 1e-6
    # K2: union
    crp, cidx, cw = oracle.combine_neighbor_sets(rp, idx, w_o, 1.0)
    g.combine(1.0)
    nl = g.export()
    assert np.array_equal(nl.row_ptr, crp) and np.array_equal(nl.idx, cidx)
    assert _rel(nl.val, cw) <= 1e-6
    E = int(crp[-1])
    print(f"1M: union {E} entries bit-exact in structure", flush=True)
    del w_o, cw, crp, cidx

    # K3: schedule of the GPU's graph
    ep = oracle.similarities_to_epochs(nl.row_ptr, nl.idx, nl.val, T, NSR)
    lay = umx.Layout(graph=g, num_epochs=T, negative_sample_rate=NSR)
    ed = lay.epoch_data()
    assert np.array_equal(ed.cumulative_num_edges, ep["row_ptr"])
    assert np.array_equal(ed.edge_targets, ep["targets"])
    assert np.array_equal(ed.epochs_per_sample, ep["eps"])
    assert np.array_equal(ed.epoch_of_next_negative_sample, ep["next_neg"])

    # K4: a captured block of 10 epochs, then a traced epoch
    a, b = umx.find_ab(1.0, 0.1)
    emb0 = np.random.default_rng(103).uniform(-10, 10, size=(n, 2)).astype(np.float32)
    lay.set_params(2, a, b, 1.0, 1.0, SEED)
    lay.set_embedding(emb0)
    lay.set_graph_block(10)
    begin, _, _, _ = lay.work_items()
    assert (begin.size + 63) // 64 > 4096  # the automatic tile is 16 columns
    assert lay.launch_shape()[0] == 16
    lay.run(10)
    orc = oracle.Layout(ep, T, NSR, rng_kind=oracle.RNG_COUNTER, seed=SEED, dtype=np.float32)
    eo = emb0.copy()
    t0 = time.perf_counter()
    orc.run(eo, a, b, epoch_limit=10)
    print(f"1M: oracle 10 epochs in {time.perf_counter() - t0:.1f}s", flush=True)
    ed = lay.epoch_data()
    nx, nn, cur, tal = orc.state()
    assert ed.current_epoch == cur == 10
    assert np.array_equal(ed.epoch_of_next_sample, nx)
    assert np.array_equal(ed.epoch_of_next_negative_sample, nn)
    assert np.array_equal(lay.tallies()[:10], tal[:10])
    del nx, nn, eo

    before = lay.get_embedding()
    cap = 2_000_000
    lay.set_trace(cap)
    lay.run(11)
    tr, tph = lay.get_trace(partial=True, with_phase=True)
    lay.set_trace(0)
    assert tr.shape[0] == cap
    kinds = tr[:, 0].astype(np.int64)
    heads = tr[:, 1].astype(np.int64)
    others = tr[:, 2].astype(np.int64)
    fwd = kinds != 2
    P, _ = lay.phases()
    after = lay.get_embedding()
    fresh = phase_of(others, P) < tph
    read = np.where(fresh[:, None], after[others], before[others])
    # a point whose row is cut into pieces is final after the last phase
    # that runs one of them; read earlier than that, it is not compared
    _, _, irow, _, iph, _ = lay.work_items(with_phase=True)
    last = phase_of(np.arange(after.shape[0]), P)
    np.maximum.at(last, irow, iph)
    pinned = ~fresh | (last[others] < tph)
    assert np.mean(pinned) > 0.9
    fwd_p = fwd & pinned
    assert np.array_equal(tr[fwd_p, 5:7].astype(np.float32), read[fwd_p])
    alpha = oracle.epoch_alpha(1.0, 10, T, np.float32)
    h = tr[fwd, 3:5].astype(np.float32)
    o = tr[fwd, 5:7].astype(np.float32)
    ref = oracle.step_delta_many(kinds[fwd].astype(np.int32), h, o, a, b, 1.0, alpha, np.float32).astype(np.float64)
    got = tr[fwd, 7:9]
    assert np.all(np.abs(got - ref) <= 1e-5 * np.abs(ref) + 1e-6)
    rev = ~fwd
    ref2 = -oracle.step_delta_many(np.zeros(int(rev.sum()), np.int32), tr[rev, 5:7].astype(np.float32),
                                   tr[rev, 3:5].astype(np.float32), a, b, 1.0, alpha, np.float32).astype(np.float64)
    assert np.all(np.abs(tr[rev, 7:9] - ref2) <= 1e-5 * np.abs(ref2) + 1e-6)
    worst = float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-6)))
    print(f"1M: {cap} traced steps of epoch 10 ({int(fwd.sum())} head / negative, {int(rev.sum())} reverse); "
          f"worst relative step error {worst:.2e}", flush=True)


def _c4_sizes():
    rng = np.random.default_rng(4)
    return np.exp(rng.uniform(np.log(2_000), np.log(200_000), size=64)).astype(np.int64)


def test_batched_64_instances(umx):
    import torch

    dev = torch.device("cuda:0")
    k, T = 15, 200
    sizes = _c4_sizes()
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    N = int(off[-1])
    idx_l, dist_l = [], []
    for q, nq in enumerate(sizes):
        xq, _ = mixture(4000 + q, int(nq), 30, 10)
        i_, d_ = umx.find_knn(xq, k)
        idx_l.append(i_.reshape(-1))
        dist_l.append(d_.reshape(-1))
    print(f"64 instances: {N} observations, kNN done", flush=True)
    seeds = np.array([SEED + q for q in range(64)], dtype=np.uint64)
    a, b = umx.find_ab(1.0, 0.1)
    emb0 = np.random.default_rng(44).uniform(-10, 10, size=(N, 2)).astype(np.float32)

    # graph per instance from the batched build vs the oracle
    g = umx.Graph.batch([umx.NeighborList(np.arange(int(nq) + 1, dtype=np.int64) * k, i_, d_)
                         for nq, i_, d_ in zip(sizes, idx_l, dist_l)])
    g.neighbor_similarities()
    g.combine(1.0)
    nl = g.export()
    parts = []
    for q, nq in enumerate(sizes):
        r0, r1 = int(off[q]), int(off[q + 1])
        e0, e1 = int(nl.row_ptr[r0]), int(nl.row_ptr[r1])
        prp, pidx, pval = nl.row_ptr[r0:r1 + 1] - e0, nl.idx[e0:e1] - r0, nl.val[e0:e1]
        rpq = np.arange(int(nq) + 1, dtype=np.int64) * k
        w_o, _, _ = oracle.neighbor_similarities(rpq, dist_l[q].astype(np.float64), nthreads=8)
        orp, oidx, oval = oracle.combine_neighbor_sets(rpq, idx_l[q], w_o, 1.0)
        assert np.array_equal(prp, orp) and np.array_equal(pidx, oidx), q
        assert _rel(pval, oval) <= 1e-6, q
        parts.append((prp, pidx, pval))
    print("64 instances: graphs equal the oracle's", flush=True)

    # schedule after 3 epochs of the batched run, per instance
    lay = umx.Layout(graph=g, num_epochs=np.full(64, T, np.int32), negative_sample_rate=NSR)
    lay.set_params(2, a, b, 1.0, 1.0, 0)
    lay.set_instance_seeds(seeds)
    lay.set_embedding(emb0)
    lay.run(3)
    ed = lay.epoch_data()
    tal = np.zeros((3, 3), np.int64)
    for q, (prp, pidx, pval) in enumerate(parts):
        r0, r1 = int(off[q]), int(off[q + 1])
        e0, e1 = int(ed.cumulative_num_edges[r0]), int(ed.cumulative_num_edges[r1])
        ep = oracle.similarities_to_epochs(prp, pidx, pval, T, NSR)
        assert np.array_equal(ed.epochs_per_sample[e0:e1], ep["eps"]), q
        orc = oracle.Layout(ep, T, NSR, rng_kind=oracle.RNG_COUNTER, seed=int(seeds[q]), dtype=np.float32)
        orc.run(emb0[r0:r1].copy(), a, b, epoch_limit=3)
        nx, nn, _, t = orc.state()
        assert np.array_equal(ed.epoch_of_next_sample[e0:e1], nx), q
        assert np.array_equal(ed.epoch_of_next_negative_sample[e0:e1], nn), q
        tal += t[:3]
    assert np.array_equal(lay.tallies()[:3], tal)
    del lay, g
    print("64 instances: schedules equal the oracle's", flush=True)

    # batched versus 64 sequential plain runs, inputs resident on the GPU
    cat_idx = torch.from_numpy(np.concatenate(idx_l)).to(dev)
    cat_dist = torch.from_numpy(np.concatenate(dist_l)).to(dev)
    rp_all = torch.arange(N + 1, dtype=torch.int64, device=dev) * k
    rps = [torch.arange(int(nq) + 1, dtype=torch.int64, device=dev) * k for nq in sizes]
    e_dev = torch.from_numpy(emb0).to(dev)
    out = torch.empty_like(e_dev)

    def plain_runs():
        for q, nq in enumerate(sizes):
            e0 = int(off[q]) * k
            gq = umx.Graph((rps[q].data_ptr(), cat_idx.data_ptr() + 4 * e0, cat_dist.data_ptr() + 4 * e0),
                           device_ptrs=True, n=int(nq), dtype=np.float32)
            gq.neighbor_similarities()
            gq.combine(1.0)
            lq = umx.Layout(graph=gq, num_epochs=T, negative_sample_rate=NSR)
            lq.set_params(2, a, b, 1.0, 1.0, int(seeds[q]))
            lq.set_embedding(device_ptr=e_dev.data_ptr() + 8 * int(off[q]))
            lq.run(T)
            lq.get_embedding(device_ptr=out.data_ptr() + 8 * int(off[q]))

    def batched_run():
        gb = umx.Graph((rp_all.data_ptr(), cat_idx.data_ptr(), cat_dist.data_ptr()), device_ptrs=True, n=N,
                       dtype=np.float32, inst_offsets=off)
        gb.neighbor_similarities()
        gb.combine(1.0)
        lb = umx.Layout(graph=gb, num_epochs=np.full(64, T, np.int32), negative_sample_rate=NSR)
        lb.set_params(2, a, b, 1.0, 1.0, 0)
        lb.set_instance_seeds(seeds)
        lb.set_embedding(device_ptr=e_dev.data_ptr())
        lb.autotune(3)
        lb.set_graph_block(25)
        lb.run(T)
        lb.get_embedding(device_ptr=out.data_ptr())

    plain_runs()
    batched_run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_runs()
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    batched_run()
    torch.cuda.synchronize()
    t_bat = time.perf_counter() - t0
    print(f"64 instances: 64 sequential plain runs {1000 * t_seq:.0f} ms, batched {1000 * t_bat:.0f} ms "
          f"({t_seq / t_bat:.2f}x)", flush=True)
    assert t_seq >= 3.0 * t_bat
