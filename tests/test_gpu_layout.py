"""GPU parity of the layout epochs against the oracle, through the C ABI.

* Schedule: after any number of epochs the production kernel's
  epoch_of_next_sample / epoch_of_next_negative_sample equal the oracle's bit
  for bit, and its per-epoch tallies (firing edges, draws, self-hits) equal
  the oracle's exactly (same counter sampler).
* Per-edge updates: both kernels record every step (head, other, coordinates
  before, move).  The step sequence (serial) or per-epoch step multiset
  (production) equals the oracle's exactly, and every step equals the
  reference's update computed from the same inputs: 1e-5 relative in f32,
  1e-12 in f64.  Whole-embedding comparisons over several epochs are not
  meaningful: the SGD is chaotic (a 1-ulp difference grows to O(1) within
  four epochs on the CPU alone).
* Final embedding: the production result and the oracle's serial result at
  the same seed and initial embedding agree on 15-NN preservation and
  trustworthiness within 0.01.
"""
import numpy as np
import pytest

import oracle
from _data import c1_knn, layout_fixture
from _quality import knn_preservation, trustworthiness
from _steps import check_pull_inputs, check_reverse_steps, check_steps, phase_of

pytestmark = pytest.mark.gpu

A, B = 1.8956, 0.8006


def _graph(nobs, k, dtype):
    x, rp, idx, w = layout_fixture(nobs, k)
    return x, rp, idx, w.astype(dtype)


def _pair(umx, rp, idx, w, T, nsr, seed, dtype):
    nl = umx.NeighborList(rp, idx, w)
    lay = umx.Layout(csr=nl, num_epochs=T, negative_sample_rate=nsr)
    ep = oracle.similarities_to_epochs(rp, idx, w, T, nsr)
    orc = oracle.Layout(ep, T, nsr, rng_kind=oracle.RNG_COUNTER, seed=seed, dtype=dtype)
    return lay, orc


# sliced-ELL launch shapes (tile columns, wavefronts per workgroup) within
# the 64 KiB LDS budget of a float workgroup
_SHAPES = [(t, w) for t in (4, 8, 16, 32) for w in (1, 2, 4) if t * w <= 64]

_SCHED = ([(nobs, k, dt, dim, 0, 1) for nobs, k in ((50, 5), (100, 10), (200, 15))
           for dt in (np.float32, np.float64) for dim in (2, 3, 5)]
          + [(200, 15, np.float32, 2, t, w) for t, w in _SHAPES])


@pytest.mark.parametrize("nobs,k,dtype,dim,tile,waves", _SCHED)
def test_schedule_and_tallies_bit_exact(umx, nobs, k, dtype, dim, tile, waves):
    x, rp, idx, w = _graph(nobs, k, dtype)
    T, nsr, seed = 120, 5.0, 77
    lay, orc = _pair(umx, rp, idx, w, T, nsr, seed, dtype)
    emb0 = np.ascontiguousarray(x[:, :dim], dtype=dtype) if dim <= 5 else None
    lay.set_params(dim, A, B, 1.0, 1.0, seed)
    lay.set_tile(tile)
    lay.set_placement(waves, 1)
    lay.set_embedding(emb0)
    lay.run(37)
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
    assert np.all(np.isfinite(emb)) and not np.array_equal(emb, emb0)


def test_reused_buffers_and_trim(umx):
    """Freed library buffers are reused by later graphs and layouts (their old
    contents must not leak into new results) and can be handed back: the
    serial replay, which is deterministic, gives the same embedding, schedule
    and tallies on fresh, reused and trimmed buffers."""
    x, rp, idx, w = _graph(200, 15, np.float32)
    T, nsr, seed = 60, 5.0, 9

    def run_once():
        g = umx.Graph(umx.NeighborList(rp, idx, w.copy()))
        g.combine(1.0)
        lay = umx.Layout(graph=g, num_epochs=T, negative_sample_rate=nsr)
        lay.set_params(2, A, B, 1.0, 1.0, seed)
        lay.set_embedding(np.ascontiguousarray(x[:, :2], dtype=np.float32))
        lay.run_serial(20)
        lay.run(T)
        ed = lay.epoch_data()
        return lay.get_embedding(), ed.epoch_of_next_sample, ed.epoch_of_next_negative_sample, lay.tallies()

    first = run_once()
    # a differently sized problem in between dirties the cached blocks
    xo, rpo, idxo, wo = _graph(100, 10, np.float32)
    lo = umx.Layout(csr=umx.NeighborList(rpo, idxo, wo), num_epochs=30, negative_sample_rate=nsr)
    lo.set_params(2, A, B, 1.0, 1.0, 3)
    lo.set_embedding(np.ascontiguousarray(xo[:, :2], dtype=np.float32))
    lo.run(30)
    del lo
    second = run_once()
    umx.trim_memory()
    third = run_once()
    for a_, b_, c_ in zip(first[1:], second[1:], third[1:]):
        assert np.array_equal(a_, b_) and np.array_equal(a_, c_)
    assert np.all(np.isfinite(first[0])) and np.all(np.isfinite(third[0]))


@pytest.mark.parametrize("dtype,rtol,atol", [(np.float64, 1e-12, 1e-15), (np.float32, 1e-5, 1e-7)])
@pytest.mark.parametrize("dim", [2, 3])
def test_serial_replay_per_step(umx, dtype, rtol, atol, dim):
    """Serial replay with the counter sampler over 25 epochs: same step
    sequence as the oracle, each step equal to the reference's update, and the
    GPU's embedding exactly the result of applying the traced steps in order."""
    x, rp, idx, w = _graph(200, 15, dtype)
    T, nsr, seed, limit = 500, 5.0, 5, 25
    lay, orc = _pair(umx, rp, idx, w, T, nsr, seed, dtype)
    emb0 = np.ascontiguousarray(x[:, :dim] * 3, dtype=dtype)
    lay.set_params(dim, A, B, 1.0, 1.0, seed)
    lay.set_embedding(emb0)
    lay.set_trace(400_000)
    lay.run_serial(limit)
    tr = lay.get_trace()
    orc.trace(400_000)
    eo = emb0.copy()
    orc.run(eo, A, B, epoch_limit=limit)
    seq = orc.traced()
    assert np.array_equal(tr[:, :3].astype(np.int64), seq)
    tal = lay.tallies()[:limit]
    counts = tal[:, 0] + tal[:, 1] - tal[:, 2]
    worst = check_steps(tr, counts, dim, dtype, A, B, 1.0, T, 0, rtol, atol)
    # sequential consistency: replaying the steps in T arithmetic reproduces
    # every traced 'before' value and the final embedding bit for bit
    cur = emb0.copy()
    for r in tr:
        kind, hi, oi = int(r[0]), int(r[1]), int(r[2])
        assert np.array_equal(cur[hi], r[3:3 + dim].astype(dtype))
        assert np.array_equal(cur[oi], r[3 + dim:3 + 2 * dim].astype(dtype))
        d = r[3 + 2 * dim:].astype(dtype)
        cur[hi] = cur[hi] + d
        if kind == 0:
            cur[oi] = cur[oi] - d
    assert np.array_equal(cur, lay.get_embedding())
    print(f"serial replay {np.dtype(dtype).name} D={dim}: {tr.shape[0]} steps, worst relative step error {worst:.2e}")


@pytest.mark.parametrize("dtype,rtol,atol", [(np.float64, 1e-12, 1e-15), (np.float32, 1e-5, 1e-7)])
def test_serial_replay_injected_mt_samples(umx, dtype, rtol, atol):
    """Config C1 graph; negative samples recorded from the oracle's
    mt19937_64 stream are injected into the GPU replay: same steps as the
    oracle, each equal to the reference's update."""
    _, (rp, idx, dist) = c1_knn()
    wd, _, _ = oracle.neighbor_similarities(rp, dist)
    crp, cidx, cw = oracle.combine_neighbor_sets(rp, idx, wd, 1.0)
    cw = cw.astype(dtype)
    T, nsr, limit = 500, 5.0, 4
    ep = oracle.similarities_to_epochs(crp, cidx, cw, T, nsr)
    orc = oracle.Layout(ep, T, nsr, rng_kind=oracle.RNG_MT, seed=1234567890, dtype=dtype)
    emb0 = np.random.default_rng(0).uniform(-10, 10, size=(2000, 2)).astype(dtype)
    eo = emb0.copy()
    orc.record(2_000_000)
    orc.trace(2_000_000)
    orc.run(eo, A, B, epoch_limit=limit)
    samples = orc.recorded()
    seq = orc.traced()
    lay = umx.Layout(csr=umx.NeighborList(crp, cidx, cw), num_epochs=T, negative_sample_rate=nsr)
    lay.set_params(2, A, B, 1.0, 1.0, 0)
    lay.set_embedding(emb0)
    lay.set_trace(seq.shape[0] + 10)
    lay.run_serial(limit, samples=samples)
    tr = lay.get_trace()
    assert np.array_equal(tr[:, :3].astype(np.int64), seq)
    tal = lay.tallies()[:limit]
    check_steps(tr, tal[:, 0] + tal[:, 1] - tal[:, 2], 2, dtype, A, B, 1.0, T, 0, rtol, atol)
    # first epochs from a shared embedding stay close despite the chaotic dynamics
    assert np.max(np.abs(lay.get_embedding() - eo)) <= 1e-3 * np.max(np.abs(eo))
    # too few samples is an error, not a fault
    lay2 = umx.Layout(csr=umx.NeighborList(crp, cidx, cw), num_epochs=T, negative_sample_rate=nsr)
    lay2.set_params(2, A, B, 1.0, 1.0, 0)
    lay2.set_embedding(emb0)
    with pytest.raises(umx.UmxError):
        lay2.run_serial(2, samples=samples[:10])


_PER_STEP = ([("pull", np.float32, dim, t, w, 32.0, 4) for dim in (2, 3) for t, w in _SHAPES]
             + [("pull", np.float32, 2, t, 1, 2.0, 4) for t in (8, 32)]
             + [("pull", np.float32, 2, 0, 1, 32.0, p) for p in (1, 2, 8)]
             + [("pull", np.float64, dim, 0, 1, 32.0, 4) for dim in (2, 3)]
             + [("push", dt, dim, 0, 1, 32.0, 4) for dt in (np.float32, np.float64) for dim in (2, 3)])


@pytest.mark.parametrize("mode,dtype,dim,tile,waves,split,phases", _PER_STEP)
def test_production_per_step(umx, mode, dtype, dim, tile, waves, split, phases):
    """Production kernels on the config C1 graph, every tile width and
    wavefronts-per-workgroup of the sliced-ELL kernel (pull mode), and a work
    split that cuts most rows into pieces.  Per epoch the steps are exactly
    the oracle's (as a multiset), every traced step equals the reference's
    update from its traced inputs within 1e-5 relative, and no step is lost:
    the embedding moves by exactly the traced steps' sum.  Push mode records
    kind 0 (head and tail move) and kind 1; pull mode also records kind 2, the
    tail move that the reverse edge gives each head.  Pull mode is
    double-buffered and runs in phases (1, 2, 4 or 8): every head starts from
    its position at the epoch's start, other points of earlier phases are
    read as this epoch left them and all others as they were at its start,
    and a reverse step is taken where the pair's head step leaves both
    points."""
    _, (rp, idx, dist) = c1_knn()
    wd, _, _ = oracle.neighbor_similarities(rp, dist)
    crp, cidx, cw = oracle.combine_neighbor_sets(rp, idx, wd, 1.0)
    cw = cw.astype(dtype)
    T, nsr, seed = 500, 5.0, 31
    lay, orc = _pair(umx, crp, cidx, cw, T, nsr, seed, dtype)
    emb = np.random.default_rng(1).uniform(-10, 10, size=(2000, dim)).astype(dtype)
    lay.set_params(dim, A, B, 1.0, 1.0, seed)
    lay.set_mode(umx.Layout.PULL if mode == "pull" else umx.Layout.PUSH)
    if mode == "pull":
        lay.set_tile(tile)
        lay.set_placement(waves, 1)
        lay.set_work_split(split)
        lay.set_phases(phases)
    lay.set_embedding(emb)
    lay.set_trace(600_000)
    orc.trace(600_000)
    eo = emb.copy()
    key = lambda a: a[np.lexsort(tuple(a[:, c] for c in range(a.shape[1] - 1, -1, -1)))]
    for limit in range(1, 7):
        before = lay.get_embedding()
        lay.run(limit)
        tr, tph = lay.get_trace(with_phase=True)
        orc.run(eo, A, B, epoch_limit=limit)
        seq = orc.traced()
        orc.trace(600_000)
        kinds = tr[:, 0].astype(np.int64)
        trip = tr[:, :3].astype(np.int64)
        assert np.array_equal(key(trip[kinds != 2]), key(seq))
        alpha = oracle.epoch_alpha(1.0, limit - 1, T, dtype)
        heads = tr[:, 1].astype(np.int64)
        others = tr[:, 2].astype(np.int64)
        if mode == "pull":
            att = seq[seq[:, 0] == 0]
            rev = trip[kinds == 2][:, [1, 2]]
            assert np.array_equal(key(rev), key(att[:, [2, 1]]))
            check_reverse_steps(tr[kinds == 2], dim, dtype, A, B, 1.0, alpha, 1e-5, 1e-6)
            P = lay.phases()[0]
            _, _, irow, isplit, iph, _ = lay.work_items(with_phase=True)
            last = phase_of(np.arange(before.shape[0]), P)
            np.maximum.at(last, irow, iph)
            share = check_pull_inputs(tr, tph, before, lay.get_embedding(), dim, dtype, P, last)
            assert tr.shape[0] == 0 or share > 0.3
            # a head's steps outside its own phase are those of a split row's
            # pieces, which run one phase after another from the row's phase
            whole = np.ones(before.shape[0], bool)
            whole[irow[isplit]] = False
            assert np.array_equal(tph[whole[heads]], phase_of(heads[whole[heads]], P))
        check_steps(tr[kinds != 2], [int((kinds != 2).sum())], dim, dtype, A, B, 1.0, T, limit - 1, 1e-5, 1e-6)
        moved = np.zeros(before.shape, np.float64)
        d = tr[:, 3 + 2 * dim:]
        np.add.at(moved, heads, d)
        if mode == "push":
            att = tr[:, 0] == 0
            np.add.at(moved, others[att], -d[att])
        after = lay.get_embedding().astype(np.float64)
        tol = (1e-4 if dtype == np.float32 else 1e-10) * max(1.0, np.max(np.abs(before)))
        assert np.max(np.abs(after - (before.astype(np.float64) + moved))) <= tol


@pytest.mark.parametrize("phases", [2, 4, 8])
def test_split_pieces_run_in_sequence(umx, phases):
    """Pull mode, config C1 graph, a work split that cuts most rows into
    pieces (threshold 8 / P expected firings).  A row cut into at most P
    pieces runs piece q in phase
    (p0 + q) mod P; its first piece of the epoch starts from the epoch's
    input, every later one from the head as the fold after the phase before
    left it: the piece's first traced step starts at the previous state plus
    the previous piece's move (its last traced position and move, minus its
    start), added in float32 as the fold adds it; the epoch's output holds the
    last fold.  Checked over several epochs for every such row."""
    _, (rp, idx, dist) = c1_knn()
    wd, _, _ = oracle.neighbor_similarities(rp, dist)
    crp, cidx, cw = oracle.combine_neighbor_sets(rp, idx, wd, 1.0)
    lay, _ = _pair(umx, crp, cidx, cw.astype(np.float32), 500, 5.0, 7, np.float32)
    emb = np.random.default_rng(3).uniform(-10, 10, size=(2000, 2)).astype(np.float32)
    lay.set_params(2, A, B, 1.0, 1.0, 7)
    lay.set_work_split(8.0 / phases)  # most rows in 2..P pieces
    lay.set_phases(phases)
    lay.set_embedding(emb)
    lay.set_trace(600_000)
    _, _, row, _, ph, from_out = lay.work_items(with_phase=True)
    count = np.bincount(row, minlength=2000)
    staged = np.nonzero((count > 1) & (count <= phases))[0]
    assert staged.size > 100
    order = np.argsort(row, kind="stable")
    first = np.searchsorted(row[order], np.arange(2000))
    checked = 0
    for limit in range(1, 6):
        before = lay.get_embedding()
        lay.run(limit)
        tr, tph = lay.get_trace(with_phase=True)
        after = lay.get_embedding()
        heads = tr[:, 1].astype(np.int64)
        key = heads * 16 + tph
        srt = np.argsort(key, kind="stable")  # a lane's records stay in their order
        ks = key[srt]
        for i in staged:
            its = order[first[i]:first[i] + count[i]]
            pph = np.sort(ph[its])
            assert np.array_equal(np.sort(from_out[its]), np.sort(pph > pph[0]))
            state = before[i].astype(np.float32)
            for p in pph:  # the row's pieces in the epoch's order
                lo, hi = np.searchsorted(ks, [i * 16 + p, i * 16 + p + 1])
                if hi > lo:
                    recs = tr[srt[lo:hi]]
                    h0 = recs[0, 3:5].astype(np.float32)
                    assert np.array_equal(h0, state), (limit, i, p)
                    end = recs[-1, 3:5].astype(np.float32) + recs[-1, 7:9].astype(np.float32)
                    state = state + (end - h0)
                    checked += 1
            assert np.array_equal(after[i], state), (limit, i)
    assert checked > 500
    print(f"{phases} phases: {staged.size} rows cut into 2..{phases} pieces, {checked} pieces checked")


_MIX = {}


def _mixture_graph():
    """10,000 x 30 float64 Gaussian mixture (10 components), exact kNN k=15,
    similarities + union from the oracle, float32 weights."""
    if not _MIX:
        from _data import knn_lists, mixture

        x, _ = mixture(11, 10000, 30, 10, dtype=np.float64)
        rp, idx, dist = knn_lists(x, 15)
        wd, _, _ = oracle.neighbor_similarities(rp, dist)
        crp, cidx, cw = oracle.combine_neighbor_sets(rp, idx, wd, 1.0)
        _MIX.update(x=x, idx=idx, g=(crp, cidx, cw.astype(np.float32)))
    return _MIX


@pytest.mark.parametrize("mode", ["pull", "push"])
def test_quality_mixture_both_modes(umx, mode):
    """Structured data (the metrics' seed-to-seed spread on the CPU is about
    0.002): both update schemes within 0.01 of the reference path at the same
    seed and initial embedding, 500 epochs."""
    m = _mixture_graph()
    crp, cidx, cw = m["g"]
    a, b = umx.find_ab(1.0, 0.1)
    emb0 = np.random.default_rng(42).uniform(-10, 10, size=(10000, 2)).astype(np.float32)
    eo = emb0.copy()
    ep = oracle.similarities_to_epochs(crp, cidx, cw, 500, 5.0)
    oracle.Layout(ep, 500, 5.0, rng_kind=oracle.RNG_MT, seed=1234567890, dtype=np.float32).run(eo, a, b)
    lay = umx.Layout(csr=umx.NeighborList(crp, cidx, cw), num_epochs=500, negative_sample_rate=5.0)
    lay.set_params(2, a, b, 1.0, 1.0, 1234567890)
    lay.set_mode(umx.Layout.PULL if mode == "pull" else umx.Layout.PUSH)
    lay.set_embedding(emb0)
    lay.run(500)
    eg = lay.get_embedding()
    hi = m["idx"].reshape(-1, 15)
    q = np.random.default_rng(0).choice(10000, 2000, replace=False)
    p_cpu, p_gpu = knn_preservation(m["x"], eo, 15, high_idx=hi), knn_preservation(m["x"], eg, 15, high_idx=hi)
    t_cpu, t_gpu = trustworthiness(m["x"], eo, 15, queries=q), trustworthiness(m["x"], eg, 15, queries=q)
    print(f"mixture 10k {mode}: preservation cpu {p_cpu:.4f} gpu {p_gpu:.4f}; "
          f"trustworthiness cpu {t_cpu:.4f} gpu {t_gpu:.4f}")
    assert abs(p_cpu - p_gpu) <= 0.01 and abs(t_cpu - t_gpu) <= 0.01


def test_serial_replay_restart_bit_exact(umx):
    x, rp, idx, w = _graph(100, 10, np.float32)
    T = 60
    emb0 = np.ascontiguousarray(x[:, :2], dtype=np.float32)
    outs = []
    for splits in ((T,), (13, 40, T)):
        lay = umx.Layout(csr=umx.NeighborList(rp, idx, w), num_epochs=T, negative_sample_rate=5.0)
        lay.set_params(2, A, B, 1.0, 1.0, 9)
        lay.set_embedding(emb0)
        for s in splits:
            lay.run_serial(s)
        outs.append(lay.get_embedding())
    assert np.array_equal(outs[0], outs[1])


def test_status_run_and_restart_schedule(umx):
    x, rp, idx, d = None, *c1_knn()[1]
    nl = umx.NeighborList(rp, idx, d.astype(np.float32))
    opt = umx.Options(initialize_method=umx.InitializeMethod.RANDOM, num_epochs=200)
    emb = np.zeros((2000, 2), np.float32)
    st = umx.initialize(nl.copy(), 2, emb, opt)
    assert st.num_epochs() == 200 and st.epoch() == 0 and st.num_observations() == 2000
    assert np.all(np.abs(emb) <= 10) and emb.std() > 1
    e1 = emb.copy()
    st.run(e1, 50)
    assert st.epoch() == 50
    st.run(e1)
    assert st.epoch() == 200
    st2 = umx.initialize(nl.copy(), 2, np.zeros((2000, 2), np.float32), opt)
    e2 = emb.copy()
    st2.run(e2)
    a1, a2 = st.get_epoch_data(), st2.get_epoch_data()
    assert np.array_equal(a1.epoch_of_next_sample, a2.epoch_of_next_sample)
    assert np.array_equal(a1.epoch_of_next_negative_sample, a2.epoch_of_next_negative_sample)
    assert np.all(np.isfinite(e1)) and np.all(np.isfinite(e2))


@pytest.mark.parametrize("fires,cap", [(32.0, 2**31 - 1), (2.0, 2**31 - 1), (4.0, 5), (0.5, 1)])
def test_work_items_cover_every_edge_once(umx, fires, cap):
    """Work items partition each row into contiguous runs, respect the split
    threshold and edge cap, mark split rows, and run longest first."""
    _, (rp, idx, dist) = c1_knn()
    wd, _, _ = oracle.neighbor_similarities(rp, dist)
    crp, cidx, cw = oracle.combine_neighbor_sets(rp, idx, wd, 1.0)
    lay = umx.Layout(csr=umx.NeighborList(crp, cidx, cw.astype(np.float32)), num_epochs=500, negative_sample_rate=5.0)
    lay.set_work_split(fires, cap)
    b, ln, row, split, ph, from_out = lay.work_items(with_phase=True)
    P, _ = lay.phases()
    # a whole row runs in its own phase; piece q (in edge order) of a row cut
    # into c pieces runs in phase (p0 + q) mod P, and starts from the head as
    # an earlier phase left it when a piece of its row ran in an earlier phase
    order = np.lexsort((b, row))
    rs, ps = row[order], ph[order]
    first = np.r_[True, rs[1:] != rs[:-1]]
    q = np.arange(rs.size) - np.maximum.accumulate(np.where(first, np.arange(rs.size), 0))
    c = np.bincount(rs)[rs]
    p0 = phase_of(rs, P)
    assert np.array_equal(ps, (p0 + q) % P)
    earliest = np.where(p0 + c - 1 >= P, 0, p0)
    assert np.array_equal(from_out[order], ps > earliest)
    ed = lay.epoch_data()
    rpe = ed.cumulative_num_edges
    assert np.all(ln >= 1) and np.all(ln <= cap)
    cover = np.zeros(rpe[-1], np.int64)
    for s_, l_ in zip(b, ln):
        cover[s_:s_ + l_] += 1
    assert np.all(cover == 1)
    assert np.all((b >= rpe[row]) & (b + ln <= rpe[row + 1]))
    nitems = np.bincount(row, minlength=2000)
    assert np.array_equal(split, nitems[row] > 1)
    assert np.array_equal(nitems > 0, np.diff(rpe) > 0)
    inv = 1.0 / ed.epochs_per_sample.astype(np.float64)
    # longest first within each phase: lengths up to 32 in exact order; a
    # longer item may be followed by one at most 1/8 longer (same length
    # class), never more
    for p in range(P):
        lp = ln[ph == p]
        pm = np.minimum.accumulate(lp)[:-1]
        nxt = lp[1:]
        assert np.all(np.where(pm <= 32, nxt <= pm, nxt * 8 < pm * 9)), "items run longest first"
        assert lp[0] >= np.max(lp) * 8 / 9
    rowf = np.add.reduceat(inv, rpe[:-1])[nitems > 1]
    assert np.all(rowf > fires) or cap < 2**31 - 1


def test_run_argument_errors(umx):
    x, rp, idx, w = _graph(50, 5, np.float32)
    lay = umx.Layout(csr=umx.NeighborList(rp, idx, w), num_epochs=10, negative_sample_rate=5.0)
    with pytest.raises(umx.UmxError):
        lay.run(5)  # no parameters / embedding yet
    lay.set_params(2, A, B)
    lay.set_embedding(np.zeros((50, 2), np.float32))
    with pytest.raises(umx.UmxError):
        lay.run(11)
    lay.run(4)
    with pytest.raises(umx.UmxError):
        lay.run(3)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_final_quality_c1_matches_reference_path(umx, dtype):
    """Config C1 (2000 x 10, k=15, 500 epochs, min_dist 0.1 -> a, b from
    find_ab), production (pull) kernel vs the reference path (serial loop,
    mt19937_64 sampler) from the same initial embedding.  C1 is unstructured
    N(0,1) data whose trustworthiness moves by about 0.008 (standard
    deviation) between seeds of the reference path itself, so both sides are
    averaged over six seeds (both are deterministic: the outcome does not
    change between runs)."""
    xh, (rp, idx, dist) = c1_knn()
    a, b = umx.find_ab(1.0, 0.1)
    T, nsr = 500, 5.0
    wd, _, _ = oracle.neighbor_similarities(rp, dist)
    crp, cidx, cw = oracle.combine_neighbor_sets(rp, idx, wd, 1.0)
    cw = cw.astype(dtype)
    emb0 = np.random.default_rng(42).uniform(-10, 10, size=(2000, 2)).astype(dtype)
    ep = oracle.similarities_to_epochs(crp, cidx, cw, T, nsr)
    hi = idx.reshape(2000, 15)
    cpu, gpu = [], []
    for seed in (1234567890, 11, 12, 13, 14, 15):
        eo = emb0.copy()
        oracle.Layout(ep, T, nsr, rng_kind=oracle.RNG_MT, seed=seed, dtype=dtype).run(eo, a, b)
        cpu.append((knn_preservation(xh, eo, 15, high_idx=hi), trustworthiness(xh, eo, 15)))
        lay = umx.Layout(csr=umx.NeighborList(crp, cidx, cw), num_epochs=T, negative_sample_rate=nsr)
        lay.set_params(2, a, b, 1.0, 1.0, seed)
        lay.set_embedding(emb0)
        lay.run(T)
        eg = lay.get_embedding()
        gpu.append((knn_preservation(xh, eg, 15, high_idx=hi), trustworthiness(xh, eg, 15)))
    (p_cpu, t_cpu), (p_gpu, t_gpu) = np.mean(cpu, axis=0), np.mean(gpu, axis=0)
    print(f"C1 {np.dtype(dtype).name} (6 seeds): 15-NN preservation cpu {p_cpu:.4f} gpu {p_gpu:.4f}; "
          f"trustworthiness cpu {t_cpu:.4f} gpu {t_gpu:.4f}")
    assert abs(p_cpu - p_gpu) <= 0.01
    assert abs(t_cpu - t_gpu) <= 0.01


@pytest.mark.parametrize("n,dtype,tile", [(20_000, np.float32, 32), (300_000, np.float32, 16),
                                          (300_000, np.float64, 16), (600_000, np.float64, 8)])
def test_automatic_tile_choice(umx, n, dtype, tile):
    """The automatic LDS tile of the production kernel: 32 columns up to 4096
    64-item chunks, then 16 for float32; for float64 16 up to 8192 chunks and
    8 above.  Ring graphs (each point's 30 neighbours at offsets +-1..15,
    weight 1 / (1 + |offset|), symmetric) keep every row one item, so the
    chunk count is about n / 64, far from either threshold."""
    offs = np.concatenate([np.arange(-15, 0), np.arange(1, 16)])
    base = 1.0 / (1.0 + np.abs(offs))
    cols = (np.arange(n, dtype=np.int64)[:, None] + offs[None, :]) % n
    order = np.argsort(cols, axis=1)
    cols = np.take_along_axis(cols, order, axis=1).astype(np.int32)
    w = np.take_along_axis(np.broadcast_to(base, cols.shape), order, axis=1).astype(dtype)
    rp = np.arange(n + 1, dtype=np.int64) * offs.size
    lay = umx.Layout(csr=umx.NeighborList(rp, cols.ravel(), w.ravel()), num_epochs=200, negative_sample_rate=5.0)
    lay.set_params(2, A, B, 1.0, 1.0, 1)
    got, _, _, chunks, _ = lay.launch_shape()
    lo, hi = {32: (0, 4096), 16: (4096, 8192 if dtype == np.float64 else 1 << 40), 8: (8192, 1 << 40)}[tile]
    assert lo < chunks <= hi, (chunks, lo, hi)
    assert got == tile
