"""Reproducibility of the production layout (GPU, through the C ABI).

The reference guarantees that a run's embedding does not depend on how it is
split into calls or on the thread count (tests/src/optimize_layout.cpp:89-125,
tests/src/umappp.cpp:104-138).  The production kernels here are
double-buffered (each epoch, phase after phase, reads the state at its start
or as earlier phases left it, and writes a new one) and fold the pieces of
split rows in edge order, so the same holds on the GPU, bit for bit, for a
given number of phases:

* repeated runs, runs stopped and restarted at any epoch, runs replayed from
  captured epoch blocks, and every launch shape (LDS tile, wavefronts per
  workgroup, placement groups) give the same embedding;
* a work split that cuts most rows into pieces is equally reproducible;
* more than four dimensions (the CSR pull kernel) too;
* a batched instance equals its plain run, bit for bit;
* the drop-in Status: run(e, 37) then run(e) equals run(e).
"""
import numpy as np
import pytest

import oracle
from _data import c1_knn, knn_lists, normal_data

pytestmark = pytest.mark.gpu

A, B = 1.8956, 0.8006
T = 120


def _c1_graph(dtype=np.float32):
    _, (rp, idx, dist) = c1_knn()
    wd, _, _ = oracle.neighbor_similarities(rp, dist)
    crp, cidx, cw = oracle.combine_neighbor_sets(rp, idx, wd, 1.0)
    return crp, cidx, cw.astype(dtype)


def _run(umx, g, dim=2, dtype=np.float32, tile=0, waves=1, split=32.0, block=0, stops=(T,), seed=4242, phases=4):
    crp, cidx, cw = g
    n = crp.size - 1
    lay = umx.Layout(csr=umx.NeighborList(crp, cidx, cw), num_epochs=T, negative_sample_rate=5.0)
    lay.set_params(dim, A, B, 1.0, 1.0, seed)
    lay.set_tile(tile)
    lay.set_placement(waves, 1)
    lay.set_work_split(split)
    lay.set_phases(phases)
    lay.set_graph_block(block)
    lay.set_embedding(np.random.default_rng(5).uniform(-10, 10, size=(n, dim)).astype(dtype))
    for s in stops:
        lay.run(s)
    ed = lay.epoch_data()
    return lay.get_embedding(), ed.epoch_of_next_sample, lay.tallies()


@pytest.mark.parametrize("split,phases", [(32.0, 4), (2.0, 4), (32.0, 1), (32.0, 8)])
def test_every_launch_shape_gives_the_same_embedding(umx, split, phases):
    g = _c1_graph()
    base = _run(umx, g, split=split, phases=phases)
    assert np.all(np.isfinite(base[0]))
    variants = [dict(tile=t, waves=w) for t in (4, 8, 16, 32) for w in (1, 2, 4) if t * w <= 64]
    variants += [dict(block=25), dict(block=64), dict(stops=(1, 37, 38, 90, T)), dict(stops=(T,))]
    for v in variants:
        got = _run(umx, g, split=split, phases=phases, **v)
        assert np.array_equal(got[0], base[0]), v
        assert np.array_equal(got[1], base[1]) and np.array_equal(got[2], base[2]), v


def test_phases_change_the_steps_not_the_schedule(umx):
    """The number of phases decides which positions a step reads, so the
    embeddings differ; the schedule and tallies do not depend on it."""
    g = _c1_graph()
    runs = [_run(umx, g, phases=p) for p in (1, 2, 4, 8)]
    for r in runs[1:]:
        assert not np.array_equal(r[0], runs[0][0])
        assert np.array_equal(r[1], runs[0][1]) and np.array_equal(r[2], runs[0][2])


def test_split_rows_are_folded(umx):
    """With a work split of 2 expected firings most rows are cut into
    pieces; the embedding still moves, stays finite and differs from the
    unsplit run (pieces start from one snapshot of the head)."""
    g = _c1_graph()
    lay = umx.Layout(csr=umx.NeighborList(*g), num_epochs=T, negative_sample_rate=5.0)
    lay.set_work_split(2.0)
    _, _, row, split = lay.work_items()
    assert split.sum() > 1000 and np.unique(row[split]).size > 300
    a = _run(umx, g, split=2.0)[0]
    b = _run(umx, g, split=32.0)[0]
    assert np.all(np.isfinite(a)) and not np.array_equal(a, b)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("dim", [3, 6])
def test_restart_and_repeat(umx, dtype, dim):
    g = _c1_graph(dtype)
    base = _run(umx, g, dim=dim, dtype=dtype)
    for stops in ((T,), (37, T), (1, 2, 3, 60, T)):
        got = _run(umx, g, dim=dim, dtype=dtype, stops=stops)
        assert np.array_equal(got[0], base[0]), stops
    assert np.array_equal(_run(umx, g, dim=dim, dtype=dtype, split=2.0)[0],
                          _run(umx, g, dim=dim, dtype=dtype, split=2.0, stops=(50, T))[0])


def test_serial_then_production_uses_the_current_buffer(umx):
    """Switching between the serial replay (in place) and the production
    kernel (double-buffered) at odd and even epochs keeps one state."""
    g = _c1_graph()
    outs = []
    for plan in (("serial", 3), ("serial", 4)):
        lay = umx.Layout(csr=umx.NeighborList(*g), num_epochs=T, negative_sample_rate=5.0)
        lay.set_params(2, A, B, 1.0, 1.0, 8)
        lay.set_embedding(np.random.default_rng(5).uniform(-10, 10, size=(2000, 2)).astype(np.float32))
        lay.run_serial(plan[1])
        e_serial = lay.get_embedding()
        lay.run(plan[1] + 5)
        lay.run_serial(plan[1] + 6)
        lay.run(T)
        outs.append((e_serial, lay.get_embedding()))
        assert np.all(np.isfinite(outs[-1][1]))
    # the state after the serial epochs is what the production run starts from
    for plan, (e_serial, _) in zip((3, 4), outs):
        lay = umx.Layout(csr=umx.NeighborList(*g), num_epochs=T, negative_sample_rate=5.0)
        lay.set_params(2, A, B, 1.0, 1.0, 8)
        lay.set_embedding(np.random.default_rng(5).uniform(-10, 10, size=(2000, 2)).astype(np.float32))
        lay.run_serial(plan)
        assert np.array_equal(lay.get_embedding(), e_serial)
        lay.run(plan + 5)
        ref = lay.get_embedding()
        lay2 = umx.Layout(csr=umx.NeighborList(*g), num_epochs=T, negative_sample_rate=5.0)
        lay2.set_params(2, A, B, 1.0, 1.0, 8)
        lay2.set_embedding(np.random.default_rng(5).uniform(-10, 10, size=(2000, 2)).astype(np.float32))
        lay2.run_serial(plan)
        lay2.run(plan + 2)
        lay2.run(plan + 5)
        assert np.array_equal(lay2.get_embedding(), ref)


def _batch(umx, sizes, epochs, seeds, dim=2):
    inst = []
    for q, n in enumerate(sizes):
        x = normal_data(800 + q, n, 5)
        inst.append(knn_lists(x, 10))
    g = umx.Graph.batch([umx.NeighborList(rp, idx, d.astype(np.float32)) for rp, idx, d in inst])
    g.neighbor_similarities()
    g.combine(1.0)
    return g


def test_batched_instance_equals_plain_run(umx):
    sizes = [700, 64, 1500, 300, 900, 1200, 40, 2200, 500]
    epochs = np.array([60, 80, 45, 80, 70, 80, 30, 80, 55], np.int32)
    seeds = np.arange(len(sizes), dtype=np.uint64) * 1000003 + 17
    g = _batch(umx, sizes, epochs, seeds)
    off = g.instances()
    exp = g.export()
    N = int(off[-1])
    emb0 = np.random.default_rng(9).uniform(-10, 10, size=(N, 2)).astype(np.float32)

    def batched(waves, groups, tile, block):
        lay = umx.Layout(graph=g, num_epochs=epochs, negative_sample_rate=5.0)
        lay.set_params(2, A, B, 1.0, 1.0, 1)
        lay.set_instance_seeds(seeds)
        lay.set_placement(waves, groups)
        lay.set_tile(tile)
        lay.set_graph_block(block)
        lay.set_embedding(emb0)
        lay.run(int(epochs.max()))
        return lay.get_embedding()

    base = batched(1, 1, 0, 0)
    for w, gr, t, bl in ((2, 8, 8, 0), (4, 8, 16, 0), (1, 8, 32, 20), (2, 1, 4, 0)):
        assert np.array_equal(batched(w, gr, t, bl), base), (w, gr, t, bl)
    for q in range(len(sizes)):
        r0, r1 = int(off[q]), int(off[q + 1])
        e0, e1 = int(exp.row_ptr[r0]), int(exp.row_ptr[r1])
        nl = umx.NeighborList(exp.row_ptr[r0:r1 + 1] - e0, exp.idx[e0:e1] - r0, exp.val[e0:e1])
        lay = umx.Layout(csr=nl, num_epochs=int(epochs[q]), negative_sample_rate=5.0)
        lay.set_params(2, A, B, 1.0, 1.0, int(seeds[q]))
        lay.set_embedding(emb0[r0:r1])
        lay.run(int(epochs[q]))
        assert np.array_equal(lay.get_embedding(), base[r0:r1]), q


def test_status_restart_equals_single_run(umx):
    """The drop-in surface (Status.hpp:101,138): run(e, 37) then run(e)
    gives exactly the embedding of run(e)."""
    _, (rp, idx, d) = c1_knn()
    nl = umx.NeighborList(rp, idx, d.astype(np.float32))
    opt = umx.Options(initialize_method=umx.InitializeMethod.RANDOM, num_epochs=150)
    e0 = np.zeros((2000, 2), np.float32)
    st = umx.initialize(nl.copy(), 2, e0, opt)
    e1 = e0.copy()
    st.run(e1, 37)
    st.run(e1)
    st2 = umx.initialize(nl.copy(), 2, np.zeros((2000, 2), np.float32), opt)
    e2 = e0.copy()
    st2.run(e2)
    assert st.epoch() == st2.epoch() == 150
    assert np.array_equal(e1, e2)
    assert not np.array_equal(e1, e0)


def test_copied_status_runs_independently(umx):
    """The reference's Status is copyable (its members are values,
    Status.hpp:42-45): a copy taken at epoch 37 and run to the end leaves
    the original at epoch 37, and gives the original's final embedding and
    schedule bit for bit, which equal a single uninterrupted run's."""
    import copy

    _, (rp, idx, d) = c1_knn()
    nl = umx.NeighborList(rp, idx, d.astype(np.float32))
    opt = umx.Options(initialize_method=umx.InitializeMethod.RANDOM, num_epochs=150)
    e0 = np.zeros((2000, 2), np.float32)
    st = umx.initialize(nl.copy(), 2, e0, opt)
    e1 = e0.copy()
    st.run(e1, 37)
    st_copy = copy.copy(st)
    assert st_copy.epoch() == 37
    e2 = e1.copy()
    st_copy.run(e2)
    assert st.epoch() == 37 and st_copy.epoch() == 150
    st.run(e1)
    assert np.array_equal(e1, e2)
    a, b = st.get_epoch_data(), st_copy.get_epoch_data()
    assert np.array_equal(a.epoch_of_next_sample, b.epoch_of_next_sample)
    assert np.array_equal(a.epoch_of_next_negative_sample, b.epoch_of_next_negative_sample)
    st3 = umx.initialize(nl.copy(), 2, np.zeros((2000, 2), np.float32), opt)
    e3 = e0.copy()
    st3.run(e3)
    assert np.array_equal(e1, e3)
