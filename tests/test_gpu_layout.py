"""GPU parity of the layout epochs against the oracle, through the C ABI.

* Schedule: after any number of epochs the Hogwild kernel's
  epoch_of_next_sample / epoch_of_next_negative_sample equal the oracle's bit
  for bit, and its per-epoch tallies (firing edges, draws, self-hits) equal
  the oracle's exactly (same counter sampler).
* Per-edge updates: the serial replay kernel against the oracle's serial loop
  over the first epochs, with the counter sampler or with samples recorded
  from the oracle's mt19937_64 stream: coordinates within 1e-5.
* Final embedding: the Hogwild result and the oracle's serial result at the
  same seed and initial embedding agree on 15-NN preservation and
  trustworthiness within 0.01.
"""
import numpy as np
import pytest

import oracle
from _data import c1_knn, layout_fixture
from _quality import knn_preservation, trustworthiness

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


@pytest.mark.parametrize("nobs,k", [(50, 5), (100, 10), (200, 15)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("dim", [2, 3, 5])
def test_schedule_and_tallies_bit_exact(umx, nobs, k, dtype, dim):
    x, rp, idx, w = _graph(nobs, k, dtype)
    T, nsr, seed = 120, 5.0, 77
    lay, orc = _pair(umx, rp, idx, w, T, nsr, seed, dtype)
    emb0 = np.ascontiguousarray(x[:, :dim], dtype=dtype) if dim <= 5 else None
    lay.set_params(dim, A, B, 1.0, 1.0, seed)
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


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 1e-5)])
@pytest.mark.parametrize("dim", [2, 3])
def test_serial_replay_counter_sampler(umx, dtype, tol, dim):
    x, rp, idx, w = _graph(200, 15, dtype)
    T, nsr, seed = 500, 5.0, 5
    lay, orc = _pair(umx, rp, idx, w, T, nsr, seed, dtype)
    emb0 = np.ascontiguousarray(x[:, :dim] * 3, dtype=dtype)
    lay.set_params(dim, A, B, 1.0, 1.0, seed)
    lay.set_embedding(emb0)
    eo = emb0.copy()
    for limit in (2, 3, 4):  # epoch 0 never fires; compare after each later epoch
        lay.run_serial(limit)
        orc.run(eo, A, B, epoch_limit=limit)
        g = lay.get_embedding()
        assert np.max(np.abs(g - eo)) <= tol * max(1.0, np.max(np.abs(eo)))
    nx, nn, cur, tal = orc.state()
    assert np.array_equal(lay.tallies()[:4], tal)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 1e-5)])
def test_serial_replay_injected_mt_samples(umx, dtype, tol):
    _, (rp, idx, dist) = c1_knn()
    wd, _, _ = oracle.neighbor_similarities(rp, dist)
    crp, cidx, cw = oracle.combine_neighbor_sets(rp, idx, wd, 1.0)
    cw = cw.astype(dtype)
    T, nsr = 500, 5.0
    ep = oracle.similarities_to_epochs(crp, cidx, cw, T, nsr)
    orc = oracle.Layout(ep, T, nsr, rng_kind=oracle.RNG_MT, seed=1234567890, dtype=dtype)
    orc.record(2_000_000)
    rng = np.random.default_rng(0)
    emb0 = rng.uniform(-10, 10, size=(2000, 2)).astype(dtype)
    eo = emb0.copy()
    orc.run(eo, A, B, epoch_limit=2)
    samples = orc.recorded()
    lay = umx.Layout(csr=umx.NeighborList(crp, cidx, cw), num_epochs=T, negative_sample_rate=nsr)
    lay.set_params(2, A, B, 1.0, 1.0, 0)
    lay.set_embedding(emb0)
    lay.run_serial(2, samples=samples)
    g = lay.get_embedding()
    assert np.max(np.abs(g - eo)) <= tol * max(1.0, np.max(np.abs(eo)))
    # too few samples is an error, not a fault
    lay2 = umx.Layout(csr=umx.NeighborList(crp, cidx, cw), num_epochs=T, negative_sample_rate=nsr)
    lay2.set_params(2, A, B, 1.0, 1.0, 0)
    lay2.set_embedding(emb0)
    with pytest.raises(umx.UmxError):
        lay2.run_serial(2, samples=samples[:10])


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
    """Config C1 (2000 x 10, k=15, 500 epochs, min_dist 0.1 -> a, b from find_ab)."""
    xh, (rp, idx, dist) = c1_knn()
    a, b = umx.find_ab(1.0, 0.1)
    T, nsr = 500, 5.0
    wd, _, _ = oracle.neighbor_similarities(rp, dist)
    crp, cidx, cw = oracle.combine_neighbor_sets(rp, idx, wd, 1.0)
    cw = cw.astype(dtype)
    rng = np.random.default_rng(42)
    emb0 = rng.uniform(-10, 10, size=(2000, 2)).astype(dtype)

    eo = emb0.copy()
    ep = oracle.similarities_to_epochs(crp, cidx, cw, T, nsr)
    oracle.Layout(ep, T, nsr, rng_kind=oracle.RNG_MT, seed=1234567890, dtype=dtype).run(eo, a, b)

    lay = umx.Layout(csr=umx.NeighborList(crp, cidx, cw), num_epochs=T, negative_sample_rate=nsr)
    lay.set_params(2, a, b, 1.0, 1.0, 1234567890)
    lay.set_embedding(emb0)
    lay.run(T)
    eg = lay.get_embedding()

    hi = idx.reshape(2000, 15)
    p_cpu = knn_preservation(xh, eo, 15, high_idx=hi)
    p_gpu = knn_preservation(xh, eg, 15, high_idx=hi)
    t_cpu = trustworthiness(xh, eo, 15)
    t_gpu = trustworthiness(xh, eg, 15)
    print(f"C1 {np.dtype(dtype).name}: 15-NN preservation cpu {p_cpu:.4f} gpu {p_gpu:.4f}; "
          f"trustworthiness cpu {t_cpu:.4f} gpu {t_gpu:.4f}")
    assert abs(p_cpu - p_gpu) <= 0.01
    assert abs(t_cpu - t_gpu) <= 0.01
