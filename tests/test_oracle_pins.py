"""Pins the oracle's similarity solve and epoch schedule to the reference's
own steps, at the precision the GPU is held to (CPU only).

The GPU is compared with the oracle bit for bit (schedule) and to 1e-6
(similarities), so the oracle itself has to follow the reference's exact
iteration, not merely reach *a* root or *a* valid schedule.  Three kinds of
check do that here:

* rows built so that the reference's stated steps fix the answer exactly,
  with no restatement involved: a start value that is already a root gives
  sigma = largest delta; an unreachable target in bisection mode doubles sigma
  64 times (the iteration cap); a zero target halves it 64 times; a row whose
  derivative underflows takes the doubling branch in Newton mode;
* a second restatement of neighbor_similarities.hpp:89-202, written here in
  plain Python from the reference text (IEEE double, the C library's exp and
  log2 through the math module), run on seeded row families chosen so that
  every branch of the iteration is taken (Newton step, Newton candidate
  rejected on either side, both midpoints, doubling, zero derivative,
  convergence at the start value, the 64-iteration cap), compared with the
  oracle's double path bit for bit in rho, sigma and weights;
* a numpy float32 replay of the epoch schedule written from
  optimize_layout.hpp:47-82,131-179 (fire test, epochs per negative sample as
  a division, truncated negative-sample count, both accumulators), which
  needs no embedding and no sampler, compared with the oracle's next,
  next_neg and per-epoch tallies bit for bit at negative-sample rates 5, 3
  and 0.7 over 200 epochs.
"""
import itertools
import math
import sys

import numpy as np
import pytest

import oracle
from _data import layout_fixture, similarity_fixture

DMAX = sys.float_info.max


def _reference_row(d, lc=1.0, bw=1.0, mkds=1e-3, newton=True, hits=None):
    """neighbor_similarities.hpp:64-202 for one row, Float_ = double.

    Returns (weights, rho, sigma); rho / sigma are None where the reference
    returns early (all weights 1).  'hits' collects the branches taken."""
    d = [float(x) for x in d]
    k = len(d)
    if k == 0:
        return [], None, None
    raw = int(lc)  # :56 from_float truncates
    interp = lc - float(raw)  # :57
    num_zero = 0  # :72-78
    for x in d:
        if x:
            break
        num_zero += 1
    if k - num_zero <= raw:  # :80-88
        return [1.0] * k, None, None
    ci = num_zero + raw  # :89
    lower = d[ci - 1] if ci > 0 else 0.0
    upper = d[ci]
    rho = lower + interp * (upper - lower)  # :92
    active = []
    num_le_rho = float(num_zero)  # :96-104
    for q in range(num_zero, k):
        if d[q] > rho:
            active.append(d[q] - rho)
        else:
            num_le_rho += 1.0
    if not active:  # :106-112
        return [1.0] * k, rho, None
    sigma = active[-1]  # :115-121
    lo, hi = 0.0, DMAX
    target = math.log2(k + 1) * bw  # :127
    it = 0
    while it < 64:  # :129-184
        observed = num_le_rho
        deriv = 0.0
        inv = 1 / sigma
        inv2 = inv * inv
        for x in active:
            cur = math.exp(-x * inv)
            observed += cur
            deriv += x * cur * inv2
        diff = observed - target
        if abs(diff) < 1e-5:
            if hits is not None:
                hits.add("converged_at_start" if it == 0 else "converged")
            break
        if diff > 0:
            hi = sigma
        else:
            lo = sigma
        ok = False
        if newton:
            if deriv:
                alt = sigma - diff / deriv
                if lo < alt < hi:
                    sigma = alt
                    ok = True
                    if hits is not None:
                        hits.add("newton")
                elif hits is not None:
                    hits.add("newton_rejected_low" if diff > 0 else "newton_rejected_high")
            elif hits is not None:
                hits.add("zero_derivative")
        if not ok:
            if diff > 0:
                sigma += (lo - sigma) / 2
                tag = "midpoint_low"
            elif hi == DMAX:
                sigma *= 2
                tag = "doubling"
            else:
                sigma += (hi - sigma) / 2
                tag = "midpoint_high"
            if hits is not None:
                hits.add(tag)
        it += 1
    if it == 64 and hits is not None:
        hits.add("iteration_cap")
    mean = 0.0  # :187-192
    for x in d:
        mean += x
    mean /= k
    sigma = max(mkds * mean, sigma)
    inv = 1 / sigma  # :194-202
    return [math.exp(-(x - rho) * inv) if x > rho else 1.0 for x in d], rho, sigma


def _rows_to_csr(rows):
    rp = np.zeros(len(rows) + 1, dtype=np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    val = np.concatenate([np.asarray(r, dtype=np.float64) for r in rows]) if rp[-1] else np.zeros(0)
    return rp, val


def _compare_with_restatement(rows, lc, bw, mkds, newton, hits):
    rp, val = _rows_to_csr(rows)
    w, rho, sigma = oracle.neighbor_similarities(rp, val, local_connectivity=lc, bandwidth=bw,
                                                 min_k_dist_scale=mkds, use_newton=newton)
    for i, d in enumerate(rows):
        ew, erho, esig = _reference_row(d, lc, bw, mkds, newton, hits)
        assert np.array_equal(w[rp[i]:rp[i + 1]], np.asarray(ew, dtype=np.float64)), i
        if esig is not None:
            assert rho[i] == erho and sigma[i] == esig, (i, rho[i], erho, sigma[i], esig)


def _row_families(seed, count):
    """Seeded rows whose iterations take every branch between them."""
    rng = np.random.default_rng(seed)
    out = []
    for t in range(count):
        k = int(rng.integers(3, 20))
        kind = t % 4
        if kind == 0:  # ordinary: exponential distances
            d = rng.exponential(1.0, k)
        elif kind == 1:  # one far neighbour: Newton overshoots from the start value
            d = np.concatenate([1 + rng.uniform(0, 1e-3, k - 1), [1 + rng.uniform(10, 1000)]])
        elif kind == 2:  # two far neighbours
            d = np.concatenate([rng.uniform(1, 1.01, k - 2), rng.uniform(5, 500, 2)])
        else:  # six decades
            d = 10 ** rng.uniform(-3, 3, k)
        out.append(np.sort(d))
    return out


@pytest.mark.parametrize("newton", [True, False])
def test_similarities_equal_restatement_on_every_branch(newton):
    rows = _row_families(20260, 800)
    hits = set()
    for bw in (0.01, 0.05, 0.3, 1.0, 2.0):
        _compare_with_restatement(rows, 1.0, bw, 1e-3, newton, hits)
    _compare_with_restatement(rows[:200], 1.4, 1.0, 1e-3, newton, hits)  # interpolated rho
    _compare_with_restatement(rows[:200], 0.0, 1.0, 1e-3, newton, hits)  # rho = 0 (ci = 0)
    if newton:
        need = {"newton", "newton_rejected_low", "newton_rejected_high", "midpoint_low", "midpoint_high",
                "doubling", "zero_derivative", "converged", "iteration_cap"}
    else:
        need = {"midpoint_low", "midpoint_high", "doubling", "converged", "iteration_cap"}
    assert need <= hits, need - hits


@pytest.mark.parametrize("nobs,k,conn", list(itertools.product([50, 100, 200], [5, 10, 20], [0.4, 1.0, 1.3, 2.5])))
def test_reference_grid_equals_restatement(nobs, k, conn):
    rp, _, dist = similarity_fixture(nobs, k, conn)
    rows = [dist[rp[i]:rp[i + 1]] for i in range(nobs)]
    _compare_with_restatement(rows, conn, 1.0, 1e-3, True, None)


def _observed(d, rho, sigma):
    """Sum over the row of exp(-max(0, d - rho) / sigma), in the reference's
    order and arithmetic (:132-138)."""
    obs = float(sum(1 for x in d if x <= rho))
    inv = 1 / sigma
    for x in d:
        if x > rho:
            obs += math.exp(-(x - rho) * inv)
    return obs


@pytest.mark.parametrize("newton", [True, False])
def test_start_value_already_a_root(newton):
    # With bandwidth set so that the target equals the observed sum at the
    # largest delta, iteration 0 stops at once: sigma is exactly that delta.
    d = [0.5, 0.8, 1.1, 1.15, 1.6, 2.3, 2.9, 3.7]
    rho, dmax = d[0], d[-1] - d[0]
    bw = _observed(d, rho, dmax) / math.log2(len(d) + 1)
    assert abs(math.log2(len(d) + 1) * bw - _observed(d, rho, dmax)) < 1e-9
    # the smallest delta is not a root, so a different start would move
    assert abs(math.log2(len(d) + 1) * bw - _observed(d, rho, d[1] - d[0])) > 0.5
    w, r, s = oracle.neighbor_similarities(np.array([0, len(d)], np.int64), np.array(d), bandwidth=bw,
                                           use_newton=newton)
    assert r[0] == rho and s[0] == dmax
    np.testing.assert_array_equal(w, [1.0] + [math.exp(-((x - rho) * (1 / dmax))) for x in d[1:]])


def test_doubling_to_a_root_in_bisection_mode():
    # The target is the observed sum at twice the largest delta: iteration 0
    # finds diff < 0 with hi still at DBL_MAX, doubles sigma, and stops.
    d = [0.25, 0.5, 0.75, 1.0, 1.25, 2.25]
    rho, dmax = d[0], d[-1] - d[0]
    bw = _observed(d, rho, 2 * dmax) / math.log2(len(d) + 1)
    assert _observed(d, rho, dmax) < math.log2(len(d) + 1) * bw - 1e-3
    _, r, s = oracle.neighbor_similarities(np.array([0, len(d)], np.int64), np.array(d), bandwidth=bw,
                                           use_newton=False)
    assert r[0] == rho and s[0] == 2 * dmax


def test_zero_derivative_takes_the_doubling_branch():
    # Distances near 1e200: 1/sigma^2 underflows to 0, so the derivative is 0
    # and even Newton mode falls back to doubling; the target is the observed
    # sum at twice the largest delta.
    scale = 1e200
    d = [1.0 * scale, 1.5 * scale, 2.0 * scale, 3.0 * scale, 4.0 * scale]
    rho, dmax = d[0], d[-1] - d[0]
    assert (1 / dmax) * (1 / dmax) == 0.0
    bw = _observed(d, rho, 2 * dmax) / math.log2(len(d) + 1)
    _, r, s = oracle.neighbor_similarities(np.array([0, len(d)], np.int64), np.array(d), bandwidth=bw,
                                           use_newton=True)
    assert r[0] == rho and s[0] == 2 * dmax


def test_iteration_cap_unreachable_target():
    # Target above any observable sum (bandwidth 10): every iteration of the
    # bisection mode doubles sigma, 64 times in all.
    d = [0.1, 0.3, 0.35, 0.9, 1.4]
    dmax = d[-1] - d[0]
    _, _, s = oracle.neighbor_similarities(np.array([0, len(d)], np.int64), np.array(d), bandwidth=10.0,
                                           use_newton=False)
    assert s[0] == dmax * 2.0 ** 64


def test_iteration_cap_zero_target():
    # Target 0: every iteration halves sigma towards lo = 0, 64 times; the
    # floor is switched off (min_k_dist_scale 0) so the capped value shows.
    d = [0.1, 0.3, 0.35, 0.9, 1.4]
    dmax = d[-1] - d[0]
    w, _, s = oracle.neighbor_similarities(np.array([0, len(d)], np.int64), np.array(d), bandwidth=0.0,
                                           min_k_dist_scale=0.0, use_newton=False)
    assert s[0] == dmax * 2.0 ** -64
    assert w[0] == 1 and np.all(w[1:] == 0)


def test_floor_from_the_mean_distance():
    # min_k_dist_scale * mean(all distances, zeros included) replaces a
    # smaller sigma (:186-192): the zero-target row's sigma becomes the floor.
    d = [0.0, 0.1, 0.3, 0.35, 0.9, 1.4]
    _, _, s = oracle.neighbor_similarities(np.array([0, len(d)], np.int64), np.array(d), bandwidth=0.0,
                                           min_k_dist_scale=0.25, use_newton=False)
    mean = 0.0
    for x in d:
        mean += x
    assert s[0] == 0.25 * (mean / len(d))


# ---------------------------------------------------------------------------
# Epoch schedule: optimize_layout.hpp:47-82 (set-up) and :131-179 (per epoch),
# in numpy float32.  The schedule does not depend on the embedding or on the
# negative samples drawn, only on epochs_per_sample.
# ---------------------------------------------------------------------------
def _schedule_replay(rp, w, num_epochs, nsr):
    f32 = np.float32
    w = np.asarray(w, dtype=f32)
    nsr = f32(nsr)  # Options double -> Float_ (initialize.hpp:115)
    maxed = f32(0)  # :49-56
    if w.size:
        maxed = max(maxed, w.max())
    limit = maxed / f32(num_epochs)  # :63 (Float_ / int)
    keep = w >= limit  # :68
    eps = maxed / w[keep]  # :70
    nxt = eps.copy()  # :77
    nxt_neg = eps / nsr  # :78-81, e /= negative_sample_rate
    fires, draws = [], []
    for n in range(num_epochs):
        epoch = f32(n)  # :131
        on = ~(nxt > epoch)  # :141-143
        e_on = eps[on]
        epns = e_on / nsr  # :160
        num_neg = ((epoch - nxt_neg[on]) / epns).astype(np.int64)  # :161, truncation
        nxt[on] = nxt[on] + e_on  # :178
        nxt_neg[on] = nxt_neg[on] + num_neg.astype(f32) * epns  # :179, int * Float_
        fires.append(int(on.sum()))
        draws.append(int(np.maximum(num_neg, 0).sum()))
    return eps, nxt, nxt_neg, np.array(fires), np.array(draws)


@pytest.mark.parametrize("nsr", [5.0, 3.0, 0.7])
@pytest.mark.parametrize("nobs,k", [(200, 15), (500, 10)])
def test_schedule_equals_float32_replay(nobs, k, nsr):
    T = 200
    _, rp, idx, w = layout_fixture(nobs, k)
    w32 = w.astype(np.float32)
    eps, nx, nn, fires, draws = _schedule_replay(rp, w32, T, nsr)
    ep = oracle.similarities_to_epochs(rp, idx, w32, T, nsr)
    assert np.array_equal(ep["eps"], eps)
    assert np.array_equal(ep["next_neg"], eps / np.float32(nsr))
    emb = np.random.default_rng(nobs).uniform(-5, 5, size=(nobs, 2)).astype(np.float32)
    lay = oracle.Layout(ep, T, nsr, rng_kind=oracle.RNG_COUNTER, seed=99, dtype=np.float32)
    lay.run(emb, 1.58, 0.9, epoch_limit=T)
    onx, onn, cur, tal = lay.state()
    assert cur == T
    assert np.array_equal(onx, nx)
    assert np.array_equal(onn, nn)
    assert np.array_equal(tal[:, 0], fires)
    assert np.array_equal(tal[:, 1], draws)


def test_schedule_replay_is_sensitive_to_the_division():
    # The negative-sample spacing is eps / nsr (a division, :160); eps * (1 / nsr)
    # rounds differently for non-dyadic rates, and the replay tells them apart.
    _, rp, idx, w = layout_fixture(200, 15)
    w32 = w.astype(np.float32)
    eps, _, nn, _, _ = _schedule_replay(rp, w32, 200, 5.0)
    alt = eps * (np.float32(1) / np.float32(5.0))
    assert not np.array_equal(eps / np.float32(5.0), alt)
