"""Edge cases of the layout through the C ABI, against the oracle.

* Ragged graphs (empty rows, hub targets with many mirrors) after the fuzzy
  union, with rows split over several work items (a low split threshold on
  the hub graphs), chunks of the production kernel mixing lengths and
  carrying padding.  Schedule and tallies stay bit-exact after T epochs, in
  both update modes and several dimensions, and observations without edges
  never move.
* Degenerate sizes (no observations, one observation, no edges at all): the
  whole path runs, moves nothing and counts nothing.
"""
import numpy as np
import pytest

import oracle
from _data import random_ragged_graph

pytestmark = pytest.mark.gpu

A, B = 1.8956, 0.8006


@pytest.mark.parametrize("seed,n,maxk,hubs", [(1, 3000, 40, 4), (2, 800, 5, 2), (3, 2000, 90, 0)])
@pytest.mark.parametrize("mode,dim", [("pull", 2), ("pull", 3), ("push", 2)])
def test_ragged_layout_schedule_bit_exact(umx, seed, n, maxk, hubs, mode, dim):
    rp, idx, d = random_ragged_graph(seed, n, maxk, np.float32, hubs)
    w, _, _ = oracle.neighbor_similarities(rp, d)
    crp, cidx, cw = oracle.combine_neighbor_sets(rp, idx, w, 1.0)
    T, nsr, s = 60, 5.0, 11
    lay = umx.Layout(csr=umx.NeighborList(crp, cidx, cw), num_epochs=T, negative_sample_rate=nsr)
    ep = oracle.similarities_to_epochs(crp, cidx, cw, T, nsr)
    orc = oracle.Layout(ep, T, nsr, rng_kind=oracle.RNG_COUNTER, seed=s, dtype=np.float32)
    emb0 = np.random.default_rng(seed).uniform(-5, 5, size=(n, dim)).astype(np.float32)
    lay.set_params(dim, A, B, 1.0, 1.0, s)
    lay.set_mode(umx.Layout.PULL if mode == "pull" else umx.Layout.PUSH)
    if hubs:
        lay.set_work_split(4.0)  # rows expected to fire more than 4 edges per epoch become several items
    lay.set_embedding(emb0)
    lay.run(T)
    eo = emb0.copy()
    orc.run(eo, A, B)
    ed = lay.epoch_data()
    nx, nn, cur, tal = orc.state()
    assert ed.current_epoch == T == cur
    assert np.array_equal(ed.epoch_of_next_sample, nx)
    assert np.array_equal(ed.epoch_of_next_negative_sample, nn)
    assert np.array_equal(lay.tallies(), tal)
    emb = lay.get_embedding()
    assert np.all(np.isfinite(emb))
    # an observation with no edges is never a head nor a tail
    lonely = np.diff(crp) == 0
    assert np.array_equal(emb[lonely], emb0[lonely])
    if hubs:
        _, _, _, split = lay.work_items()
        assert split.any()


@pytest.mark.parametrize("n", [0, 1, 5])
def test_graph_without_edges(umx, n):
    rp = np.zeros(n + 1, np.int64)
    g = umx.Graph(umx.NeighborList(rp, np.zeros(0, np.int32), np.zeros(0, np.float32)))
    g.neighbor_similarities()
    g.combine(1.0)
    nl = g.export()
    assert nl.row_ptr.size == n + 1 and nl.idx.size == 0
    lay = umx.Layout(graph=g, num_epochs=10, negative_sample_rate=5.0)
    emb0 = np.arange(n * 2, dtype=np.float32).reshape(n, 2)
    lay.set_params(2, A, B, 1.0, 1.0, 3)
    lay.set_embedding(emb0)
    lay.run(10)
    assert np.array_equal(lay.get_embedding(), emb0)
    assert int(lay.tallies().sum()) == 0
    assert lay.epoch_data().current_epoch == 10


def test_batch_with_degenerate_and_ragged_instances(umx):
    """A batch mixing an ordinary kNN instance, a ragged one with hubs, and
    instances with 0, 1 and 5 observations and no edges: every instance's
    schedule equals the oracle's run of it alone, tallies add up, and
    observations without edges never move."""
    from test_gpu_batch import _check_schedule, _instances, _split

    (_, rp0, idx0, d0), = _instances([300], 10, 900, np.float32)
    rp1, idx1, d1 = random_ragged_graph(4, 400, 30, np.float32, 3)
    empty = lambda m: umx.NeighborList(np.zeros(m + 1, np.int64), np.zeros(0, np.int32), np.zeros(0, np.float32))
    lists = [umx.NeighborList(rp0, idx0, d0), empty(0), umx.NeighborList(rp1, idx1, d1), empty(1), empty(5)]
    g = umx.Graph.batch(lists)
    g.neighbor_similarities()
    g.combine(1.0)
    off = g.instances()
    assert np.array_equal(off, [0, 300, 300, 700, 701, 706])
    parts = _split(g.export(), off)
    epochs = [40, 10, 35, 20, 40]
    seeds = [5, 6, 7, 8, 9]
    lay = umx.Layout(graph=g, num_epochs=np.asarray(epochs, np.int32), negative_sample_rate=5.0)
    emb0 = np.random.default_rng(8).uniform(-10, 10, size=(int(off[-1]), 2)).astype(np.float32)
    lay.set_params(2, A, B, 1.0, 1.0, 1)
    lay.set_instance_seeds(seeds)
    lay.set_embedding(emb0)
    orcs = []
    for q, (rp, idx, w) in enumerate(parts):
        ep = oracle.similarities_to_epochs(rp, idx, w, epochs[q], 5.0)
        orcs.append((ep, epochs[q], oracle.Layout(ep, epochs[q], 5.0, rng_kind=oracle.RNG_COUNTER, seed=seeds[q],
                                                  dtype=np.float32)))
    lay.run(40)
    _check_schedule(lay, off, orcs, emb0, 40)
    emb = lay.get_embedding()
    ed = lay.epoch_data()
    lonely = np.diff(ed.cumulative_num_edges) == 0
    assert lonely[off[3]:].all()
    assert np.array_equal(emb[lonely], emb0[lonely]) and np.all(np.isfinite(emb))


def test_one_way_edges_run_in_push_mode(umx):
    """A graph that did not go through the fuzzy union may hold an entry
    without its reverse.  Pull mode would apply a reverse step that does not
    exist, so such a layout runs in push mode (every stored edge moves its
    head and its tail, as in the reference's loop) and refuses pull mode;
    per epoch its steps equal the oracle's; a batched layout of it is
    refused.  Combined graphs and symmetric raw CSR stay in pull mode."""
    rp, idx, d = random_ragged_graph(7, 600, 12, np.float32, 0)
    w, _, _ = oracle.neighbor_similarities(rp, d)
    crp, cidx, cw = oracle.combine_neighbor_sets(rp, idx, w, 1.0)
    sym = umx.Layout(csr=umx.NeighborList(crp, cidx, cw), num_epochs=30, negative_sample_rate=5.0)
    assert sym.mode() == umx.Layout.PULL
    # one-way: drop the reverse of one entry from a symmetric graph
    r = int(np.argmax(np.diff(crp) > 2))
    t = int(cidx[crp[r]])
    rev = int(crp[t] + np.nonzero(cidx[crp[t]:crp[t + 1]] == r)[0][0])
    keep = np.ones(cidx.size, bool)
    keep[rev] = False
    orp = crp.copy()
    orp[t + 1:] -= 1
    oidx, ow = cidx[keep], cw[keep]
    T, nsr, s = 30, 5.0, 3
    lay = umx.Layout(csr=umx.NeighborList(orp, oidx, ow), num_epochs=T, negative_sample_rate=nsr)
    assert lay.mode() == umx.Layout.PUSH
    with pytest.raises(umx.UmxError) as e:
        lay.set_mode(umx.Layout.PULL)
    assert e.value.code == -4
    ep = oracle.similarities_to_epochs(orp, oidx, ow, T, nsr)
    orc = oracle.Layout(ep, T, nsr, rng_kind=oracle.RNG_COUNTER, seed=s, dtype=np.float32)
    emb0 = np.random.default_rng(1).uniform(-5, 5, size=(600, 2)).astype(np.float32)
    lay.set_params(2, A, B, 1.0, 1.0, s)
    lay.set_embedding(emb0)
    lay.set_trace(400_000)
    orc.trace(400_000)
    eo = emb0.copy()
    key = lambda a: a[np.lexsort(tuple(a[:, c] for c in range(a.shape[1] - 1, -1, -1)))]
    for limit in (5, 12, T):
        lay.run(limit)
        orc.run(eo, A, B, epoch_limit=limit)
        assert np.array_equal(key(lay.get_trace()[:, :3].astype(np.int64)), key(orc.traced()))
        lay.set_trace(400_000)
        orc.trace(400_000)
    ed = lay.epoch_data()
    nx, nn, _, tal = orc.state()
    assert np.array_equal(ed.epoch_of_next_sample, nx) and np.array_equal(lay.tallies(), tal)
    with pytest.raises(umx.UmxError) as e:
        g = umx.Graph.batch([umx.NeighborList(orp, oidx, ow), umx.NeighborList(crp, cidx, cw)])
        umx.Layout(graph=g, num_epochs=np.array([T, T], np.int32), negative_sample_rate=nsr)
    assert e.value.code == -2
