"""Batched graphs and layouts (many independent problems in the same launches).

A batched instance must behave exactly like the same problem run alone:

* graph build: similarities and the fuzzy union of every instance equal the
  plain GPU path's bit for bit (and the oracle's, as in the plain tests);
* schedule: per-instance largest weight, survivor limit and num_epochs; after
  any number of epochs every instance's epoch_of_next_sample /
  epoch_of_next_negative_sample equal the oracle's run of that instance alone,
  bit for bit, and the per-epoch tallies equal the sum over instances;
* negative samples stay inside their instance and, keyed by the instance's
  seed and its own entry numbers, are the very samples of the plain run
  (same (kind, head, other) multiset per epoch);
* every traced step equals the reference's update from its traced inputs;
* final quality of each instance within 0.01 of its plain run.
"""
import numpy as np
import pytest

import oracle
from _data import knn_lists, mixture, normal_data
from _quality import knn_preservation, trustworthiness
from _steps import check_reverse_steps, check_steps

pytestmark = pytest.mark.gpu

A, B = 1.8956, 0.8006


def _instances(sizes, k, seed0, dtype=np.float32):
    """kNN lists of independent N(0,1) data sets, one per size."""
    out = []
    for q, n in enumerate(sizes):
        x = normal_data(seed0 + q, n, 5)
        rp, idx, dist = knn_lists(x, k)
        out.append((x, rp, idx, dist.astype(dtype)))
    return out


def _batch_graph(umx, inst):
    return umx.Graph.batch([umx.NeighborList(rp, idx, d) for _, rp, idx, d in inst])


def _split(nl, off):
    """Per-instance (row_ptr, local idx, val) of a batched CSR export."""
    out = []
    for q in range(off.size - 1):
        r0, r1 = int(off[q]), int(off[q + 1])
        rp = nl.row_ptr[r0:r1 + 1] - nl.row_ptr[r0]
        e0, e1 = int(nl.row_ptr[r0]), int(nl.row_ptr[r1])
        out.append((rp, nl.idx[e0:e1] - r0, nl.val[e0:e1]))
    return out


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batched_graph_equals_plain(umx, dtype):
    sizes = [300, 57, 1000, 16, 640]
    inst = _instances(sizes, 10, 500, dtype)
    g = _batch_graph(umx, inst)
    off = g.instances()
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(sizes)]))
    g.neighbor_similarities()
    g.combine(1.0)
    parts = _split(g.export(), off)
    for (x, rp, idx, d), (brp, bidx, bval) in zip(inst, parts):
        p = umx.Graph(umx.NeighborList(rp, idx, d))
        p.neighbor_similarities()
        p.combine(1.0)
        e = p.export()
        assert np.array_equal(brp, e.row_ptr) and np.array_equal(bidx, e.idx) and np.array_equal(bval, e.val)
        # and the oracle's double path
        w, _, _ = oracle.neighbor_similarities(rp, d.astype(np.float64))
        orp, oidx, oval = oracle.combine_neighbor_sets(rp, idx, w.astype(dtype).astype(np.float64), 1.0)
        assert np.array_equal(brp, orp) and np.array_equal(bidx, oidx)
        assert np.allclose(bval, oval, rtol=1e-6 if dtype == np.float32 else 1e-12, atol=0)


def test_batched_graph_rejects_foreign_index(umx):
    inst = _instances([40, 30], 5, 900)
    bad = [list(t) for t in inst]
    bad[1][2] = bad[1][2].copy()
    bad[1][2][7] = 35  # local index beyond instance 1's 30 observations
    with pytest.raises(umx.UmxError) as e:
        umx.Graph.batch([umx.NeighborList(rp, idx, d) for _, rp, idx, d in bad])
    assert e.value.code == -2


def _batch_setup(umx, sizes, epochs, seeds, dtype=np.float32, k=10, dim=2):
    inst = _instances(sizes, k, 700, dtype)
    g = _batch_graph(umx, inst)
    g.neighbor_similarities()
    g.combine(1.0)
    off = g.instances()
    parts = _split(g.export(), off)
    lay = umx.Layout(graph=g, num_epochs=np.asarray(epochs, np.int32), negative_sample_rate=5.0)
    emb0 = np.random.default_rng(3).uniform(-10, 10, size=(int(off[-1]), dim)).astype(dtype)
    lay.set_params(dim, A, B, 1.0, 1.0, 12345)
    lay.set_instance_seeds(seeds)
    lay.set_embedding(emb0)
    orcs = []
    for q, (rp, idx, w) in enumerate(parts):
        ep = oracle.similarities_to_epochs(rp, idx, w, int(epochs[q]), 5.0)
        orcs.append((ep, int(epochs[q]), oracle.Layout(ep, int(epochs[q]), 5.0, rng_kind=oracle.RNG_COUNTER,
                                                       seed=int(seeds[q]), dtype=dtype)))
    return lay, off, parts, orcs, emb0


def _check_schedule(lay, off, orcs, emb0, upto):
    ed = lay.epoch_data()
    tal = np.zeros_like(lay.tallies())
    for q, (ep, total, orc) in enumerate(orcs):
        r0, r1 = int(off[q]), int(off[q + 1])
        e0, e1 = int(ed.cumulative_num_edges[r0]), int(ed.cumulative_num_edges[r1])
        assert np.array_equal(ed.cumulative_num_edges[r0:r1 + 1] - e0, ep["row_ptr"])
        assert np.array_equal(ed.edge_targets[e0:e1] - r0, ep["targets"])
        assert np.array_equal(ed.epochs_per_sample[e0:e1], ep["eps"])
        eo = emb0[r0:r1].copy()
        orc.run(eo, A, B, epoch_limit=min(upto, total))
        nx, nn, cur, t = orc.state()
        assert np.array_equal(ed.epoch_of_next_sample[e0:e1], nx), q
        assert np.array_equal(ed.epoch_of_next_negative_sample[e0:e1], nn), q
        tal[:t.shape[0]] += t
    assert np.array_equal(lay.tallies(), tal)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("waves,groups", [(1, 1), (2, 3), (4, 8), (1, 0)])
def test_batched_schedule_bit_exact(umx, dtype, waves, groups):
    sizes = [400, 90, 1200, 33, 800, 250, 60, 500, 150]
    epochs = [60, 35, 60, 20, 45, 60, 10, 50, 60]
    seeds = [11 + 7 * q for q in range(len(sizes))]
    lay, off, parts, orcs, emb0 = _batch_setup(umx, sizes, epochs, seeds, dtype)
    lay.set_placement(waves, groups)
    assert np.array_equal(lay.instances()[1], epochs)
    lay.run(17)
    lay.run(60)
    _check_schedule(lay, off, orcs, emb0, 60)
    # every real entry is covered by exactly one work item; padding items are empty
    b, ln, row, _ = lay.work_items()
    nnz = lay.info()[1]
    cover = np.zeros(nnz, np.int64)
    for s, l in zip(b[ln > 0], ln[ln > 0]):
        cover[s:s + l] += 1
    assert np.all(cover == 1)
    assert np.all(np.isfinite(lay.get_embedding()))


def test_batched_graph_blocks(umx):
    """Epochs replayed from captured HIP graphs (the kernel advances a device
    epoch counter) give the same schedule as direct launches."""
    sizes = [300, 700, 120, 900]
    epochs = [40, 40, 25, 40]
    seeds = [5, 6, 7, 8]
    lay, off, parts, orcs, emb0 = _batch_setup(umx, sizes, epochs, seeds)
    lay.set_graph_block(7)
    lay.run(3)
    lay.run(31)   # 4 replays of one graph + direct launches
    lay.run(40)   # one more replay, then the rest
    _check_schedule(lay, off, orcs, emb0, 40)


def test_batched_graph_blocks_longer_than_tally_ring(umx):
    """A captured block of 70 epochs folds its tallies in two parts (64 + 6
    epochs, the partial-tally ring holding 64); schedule and per-epoch
    tallies equal the oracle's, across replays and the direct launches
    after them."""
    sizes = [300, 500, 150]
    epochs = [160, 120, 160]
    seeds = [21, 22, 23]
    lay, off, parts, orcs, emb0 = _batch_setup(umx, sizes, epochs, seeds)
    lay.set_graph_block(70)
    lay.run(5)
    lay.run(160)  # two replays of one graph, then 15 direct launches
    _check_schedule(lay, off, orcs, emb0, 160)


def test_batched_samples_match_plain_runs(umx):
    """Per epoch, each instance's traced (kind, head, other) steps equal those
    of the same problem run alone with its seed; negatives stay inside their
    instance; each step equals the reference's update."""
    sizes = [500, 260, 900]
    epochs = [30, 30, 30]
    seeds = [101, 202, 303]
    lay, off, parts, orcs, emb0 = _batch_setup(umx, sizes, epochs, seeds, dim=2)
    lay.set_trace(2_000_000)
    plains = []
    for q, (rp, idx, w) in enumerate(parts):
        p = umx.Layout(csr=umx.NeighborList(rp, idx, w), num_epochs=epochs[q], negative_sample_rate=5.0)
        p.set_params(2, A, B, 1.0, 1.0, seeds[q])
        p.set_embedding(emb0[off[q]:off[q + 1]])
        p.set_trace(2_000_000)
        plains.append(p)
    key = lambda a: a[np.lexsort(tuple(a[:, c] for c in range(a.shape[1] - 1, -1, -1)))]
    inst_of = np.repeat(np.arange(len(sizes)), sizes)
    for limit in range(1, 9):
        lay.run(limit)
        tr = lay.get_trace()
        trip = tr[:, :3].astype(np.int64)
        assert np.array_equal(inst_of[trip[:, 1]], inst_of[trip[:, 2]])  # samples stay inside the instance
        for q, p in enumerate(plains):
            p.run(limit)
            pt = p.get_trace()[:, :3].astype(np.int64)
            mine = trip[inst_of[trip[:, 1]] == q].copy()
            mine[:, 1:] -= off[q]
            assert np.array_equal(key(mine), key(pt)), (limit, q)
        alpha = oracle.epoch_alpha(1.0, limit - 1, 30, np.float32)
        kinds = tr[:, 0].astype(np.int64)
        check_steps(tr[kinds != 2], [int((kinds != 2).sum())], 2, np.float32, A, B, 1.0, 30, limit - 1, 1e-5, 1e-6)
        check_reverse_steps(tr[kinds == 2], 2, np.float32, A, B, 1.0, alpha, 1e-5, 1e-6)


def test_batched_autotune_restores_state(umx):
    sizes = [600, 1500, 300, 900, 2000, 450]
    epochs = [50] * 6
    seeds = list(range(6))
    lay, off, parts, orcs, emb0 = _batch_setup(umx, sizes, epochs, seeds)
    before = lay.epoch_data()
    tile, waves, groups = lay.autotune(2)
    assert tile in (4, 8, 16, 32) and waves in (1, 2, 4) and groups >= 1
    after = lay.epoch_data()
    assert after.current_epoch == 0
    assert np.array_equal(after.epoch_of_next_sample, before.epoch_of_next_sample)
    assert np.array_equal(after.epoch_of_next_negative_sample, before.epoch_of_next_negative_sample)
    assert np.array_equal(lay.get_embedding(), emb0)
    assert not lay.tallies().any()
    # a second handle of the same shape takes the cached choice (applied
    # before its work items are built) and runs bit-exact with it
    lay2, off2, _, orcs2, emb02 = _batch_setup(umx, sizes, epochs, seeds)
    assert lay2.autotune(2) == (tile, waves, groups)
    lay.run(50)
    _check_schedule(lay, off, orcs, emb0, 50)
    lay2.run(50)
    _check_schedule(lay2, off2, orcs2, emb02, 50)


def test_batched_layout_restrictions(umx):
    lay, *_ = _batch_setup(umx, [100, 120], [20, 20], [1, 2])
    lay.set_mode(umx.Layout.PUSH)
    with pytest.raises(umx.UmxError) as e:
        lay.run(5)
    assert e.value.code == -4
    lay.set_mode(umx.Layout.PULL)
    with pytest.raises(umx.UmxError) as e:
        lay.run_serial(2)
    assert e.value.code == -4
    lay.run(20)


def test_batched_quality_matches_plain(umx):
    """Four mixture instances (1,500-6,000 points, 200 epochs): each batched
    instance within 0.01 of its plain GPU run on 15-NN preservation and
    trustworthiness, averaged over the instances."""
    sizes = [1500, 6000, 2500, 4000]
    inst, hi = [], []
    for q, n in enumerate(sizes):
        x, _ = mixture(60 + q, n, 30, 10, dtype=np.float64)
        rp, idx, dist = knn_lists(x, 15)
        inst.append((x, rp, idx, dist.astype(np.float32)))
        hi.append(idx.reshape(-1, 15))
    g = _batch_graph(umx, inst)
    g.neighbor_similarities()
    g.combine(1.0)
    off = g.instances()
    parts = _split(g.export(), off)
    a, b = umx.find_ab(1.0, 0.1)
    emb0 = np.random.default_rng(9).uniform(-10, 10, size=(int(off[-1]), 2)).astype(np.float32)
    seeds = [1000 + q for q in range(len(sizes))]
    lay = umx.Layout(graph=g, num_epochs=np.full(len(sizes), 200, np.int32), negative_sample_rate=5.0)
    lay.set_params(2, a, b, 1.0, 1.0, 0)
    lay.set_instance_seeds(seeds)
    lay.set_embedding(emb0)
    lay.run(200)
    eb = lay.get_embedding()
    pres_b, pres_p, tw_b, tw_p = [], [], [], []
    for q, (rp, idx, w) in enumerate(parts):
        p = umx.Layout(csr=umx.NeighborList(rp, idx, w), num_epochs=200, negative_sample_rate=5.0)
        p.set_params(2, a, b, 1.0, 1.0, seeds[q])
        p.set_embedding(emb0[off[q]:off[q + 1]])
        p.run(200)
        ep_ = p.get_embedding()
        mine = eb[off[q]:off[q + 1]]
        x = inst[q][0]
        pres_b.append(knn_preservation(x, mine, 15, high_idx=hi[q]))
        pres_p.append(knn_preservation(x, ep_, 15, high_idx=hi[q]))
        tw_b.append(trustworthiness(x, mine, 15))
        tw_p.append(trustworthiness(x, ep_, 15))
    print(f"batched vs plain: preservation {np.mean(pres_b):.4f} vs {np.mean(pres_p):.4f}, "
          f"trustworthiness {np.mean(tw_b):.4f} vs {np.mean(tw_p):.4f}")
    assert abs(np.mean(pres_b) - np.mean(pres_p)) <= 0.01
    assert abs(np.mean(tw_b) - np.mean(tw_p)) <= 0.01
