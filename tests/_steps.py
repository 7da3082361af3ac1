"""Per-step checks of traced layout updates against the oracle's step
function (the reference's update, optimize_layout.hpp:145-176)."""
import numpy as np

import oracle


def check_steps(trace, counts, dim, dtype, a, b, gamma, T, first_epoch, rtol, atol):
    """Every traced step's head move equals the oracle's step function applied
    to the traced coordinates before the step; returns the worst error."""
    worst = 0.0
    pos = 0
    for e, m in enumerate(counts):
        m = int(m)
        alpha = oracle.epoch_alpha(1.0, first_epoch + e, T, dtype)
        r = trace[pos:pos + m]
        h = r[:, 3:3 + dim].astype(dtype)
        o = r[:, 3 + dim:3 + 2 * dim].astype(dtype)
        ref = oracle.step_delta_many(r[:, 0].astype(np.int32), h, o, a, b, gamma, alpha, dtype).astype(np.float64)
        got = r[:, 3 + 2 * dim:3 + 3 * dim]
        err = np.abs(got - ref)
        bad = np.any(err > rtol * np.abs(ref) + atol, axis=1)
        assert not bad.any(), (e, r[bad][:3, :3], got[bad][:3], ref[bad][:3])
        if m:
            worst = max(worst, float(np.max(err / np.maximum(np.abs(ref), 1e-30))))
        pos += m
    assert pos == trace.shape[0]
    return worst


def phase_of(index, phases):
    """The phase of observations with these (instance-local) indices: blocks
    of 16 consecutive indices, block b in phase b modulo the number of phases
    (a power of two)."""
    return ((np.asarray(index, dtype=np.int64) >> 4) & (phases - 1)).astype(np.int64)


def check_pull_inputs(trace, launch_phase, before, after, dim, dtype, phases, last_phase=None):
    """One pull-mode epoch's reads: every head's first step starts from its
    position at the epoch's start; a head step (kind 0) or negative sample
    (kind 1) reads the other point as this epoch left it if that point is of
    an earlier phase than the launch that took the step, else as it was at
    the epoch's start ('before'); a reverse step (kind 2) directly follows
    the head step g of the same edge (kind 0): the head from y_i + g, the
    other from y_t - g (where the pair's first event leaves the tail).  A
    point whose row is cut into pieces moves in every phase that runs one of
    its pieces (last_phase: the last such phase of each point; whole rows:
    their own phase): read by a later launch than its last phase it is at
    its final position ('after'); read in between, at a state of the epoch
    the trace does not pin, it is not compared."""
    kinds = trace[:, 0].astype(np.int64)
    heads = trace[:, 1].astype(np.int64)
    others = trace[:, 2].astype(np.int64)
    h = trace[:, 3:3 + dim].astype(dtype)
    o = trace[:, 3 + dim:3 + 2 * dim].astype(dtype)
    g = trace[:, 3 + 2 * dim:].astype(dtype)
    launch_phase = np.asarray(launch_phase, dtype=np.int64)
    fresh = phase_of(others, phases) < launch_phase
    read = np.where(fresh[:, None], after[others], before[others])
    last = phase_of(others, phases) if last_phase is None else np.asarray(last_phase)[others]
    pinned = ~fresh | (last < launch_phase)
    fwd = (kinds != 2) & pinned
    assert np.array_equal(o[fwd], read[fwd])
    _, first = np.unique(heads, return_index=True)
    assert np.array_equal(h[first], before[heads[first]])
    n = before.shape[0]
    k0 = np.nonzero(kinds == 0)[0]
    k2 = np.nonzero(kinds == 2)[0]
    key0 = heads[k0] * n + others[k0]
    order = np.argsort(key0, kind="stable")
    key0 = key0[order]
    assert np.all(np.diff(key0) > 0)  # one head step per edge and epoch
    at = np.searchsorted(key0, heads[k2] * n + others[k2])
    assert np.all(at < key0.size)
    m0 = k0[order[at]]
    assert np.array_equal(key0[at], heads[k2] * n + others[k2])
    assert np.array_equal(h[k2], h[m0] + g[m0])
    p2 = pinned[k2]
    assert np.array_equal(o[k2][p2], (read[k2] - g[m0])[p2])
    return float(np.mean(pinned)) if pinned.size else 1.0


def check_reverse_steps(trace, dim, dtype, a, b, gamma, alpha, rtol, atol):
    """Kind-2 records (pull mode): head i's move is the tail move of edge
    t->i, minus the head move of t evaluated at the traced (y_t, y_i); every
    record is checked."""
    h = trace[:, 3:3 + dim].astype(dtype)
    o = trace[:, 3 + dim:3 + 2 * dim].astype(dtype)
    ref = -oracle.step_delta_many(np.zeros(len(trace), np.int32), o, h, a, b, gamma, alpha, dtype).astype(np.float64)
    got = trace[:, 3 + 2 * dim:]
    bad = np.any(np.abs(got - ref) > rtol * np.abs(ref) + atol, axis=1)
    assert not bad.any(), (trace[bad][:3, :3], got[bad][:3], ref[bad][:3])
