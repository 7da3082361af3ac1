"""Pins the oracle's UMAP cross-entropy loss (the autograd loss's checker).

* the closed-form gradient (embedding, a, b) equals central finite
  differences of the loss value;
* the incidence-list finite differences used at full size equal finite
  differences of the whole loss;
* per term, minus the gradient with respect to the head is the reference's
  update direction (optimize_layout.hpp:145-176): exactly for the attractive
  term, and for the repulsive term up to the factor s / (0.001 + s) of the
  reference's regulariser;
* negative samples are the counter sampler's draws (the layout's), and a
  draw equal to the head is skipped.
"""
import numpy as np
import pytest

import oracle

A, B, GAMMA = 1.5769, 0.8951, 1.7


def _problem(seed=0, n=60, dim=3, deg=5, m=5, coef=True, step=3):
    rng = np.random.default_rng(seed)
    emb = rng.uniform(-3, 3, size=(n, dim))
    heads, tails = [], []
    for i in range(n):
        for j in rng.choice(np.delete(np.arange(n), i), size=deg, replace=False):
            heads.append(i)
            tails.append(int(j))
    heads, tails = np.array(heads), np.array(tails)
    c = rng.uniform(0.2, 1.0, size=heads.size) if coef else None
    nn = rng.integers(0, m + 1, size=heads.size) if coef else None
    keys = rng.permutation(heads.size * 3)[:heads.size]
    return oracle.LossProblem(emb, heads, tails, A, B, GAMMA, m=m, seed=11, step=step, keys=keys, coef=c,
                              num_neg=nn)


@pytest.mark.parametrize("coef", [True, False])
def test_gradient_matches_finite_differences(coef):
    P = _problem(coef=coef)
    g, gab = P.grad()
    h = 1e-6
    worst = 0.0
    for i in range(P.n):
        for d in range(P.dim):
            e = P.emb.copy()
            e[i, d] += h
            up = P.loss(e)
            e[i, d] -= 2 * h
            dn = P.loss(e)
            fd = (up - dn) / (2 * h)
            assert abs(fd - g[i, d]) <= 1e-6 * max(1.0, abs(fd)), (i, d, fd, g[i, d])
            worst = max(worst, abs(fd - g[i, d]))
    for which in range(2):
        h = 1e-6
        ab = [P.a, P.b]
        ab[which] += h
        up = P.loss(None, *ab)
        ab[which] -= 2 * h
        dn = P.loss(None, *ab)
        fd = (up - dn) / (2 * h)
        assert abs(fd - gab[which]) <= 1e-6 * max(1.0, abs(fd))


def test_incidence_fd_equals_whole_loss_fd():
    P = _problem(seed=4)
    pts = np.arange(P.n)
    dims = np.arange(P.n) % P.dim
    local = P.fd(pts, dims, h=1e-6)
    for q in range(P.n):
        e = P.emb.copy()
        e[pts[q], dims[q]] += 1e-6
        up = P.loss(e)
        e[pts[q], dims[q]] -= 2e-6
        dn = P.loss(e)
        assert abs((up - dn) / 2e-6 - local[q]) <= 1e-6 * max(1.0, abs(local[q]))


def _one_term(kind, yi, yj, gamma=GAMMA):
    """Loss of a single term: one edge 0 -> 1 (attractive) or one negative
    sample of edge 0 -> 2 that lands on observation 1 (repulsive)."""
    if kind == 0:
        emb = np.array([yi, yj])
        return oracle.LossProblem(emb, [0], [1], A, B, gamma, m=0)
    # find a key whose first draw over 3 observations is observation 1
    emb = np.array([yi, yj, np.asarray(yi) + 1e6])  # the edge's tail is far away
    for key in range(1000):
        if oracle.counter_draw(0, 0, key, 0, 3) == 1:
            return oracle.LossProblem(emb, [0], [2], A, B, gamma, m=1, keys=[key])
    raise AssertionError("no key found")


@pytest.mark.parametrize("dim", [2, 3])
def test_term_gradients_match_reference_update(dim):
    rng = np.random.default_rng(dim)
    for _ in range(20):
        yi = rng.uniform(-2, 2, size=dim)
        yj = yi + rng.uniform(-3, 3, size=dim)
        s = float(np.sum((yi - yj) ** 2))
        # attractive: the reference's step at alpha = 1 is minus the gradient (no clip here)
        g, _ = _one_term(0, yi, yj).grad()
        step = oracle.step_delta(0, yi, yj, A, B, GAMMA, 1.0, np.float64)
        if np.all(np.abs(step) < 4):
            assert np.allclose(-g[0], step, rtol=1e-12, atol=1e-15)
            assert np.allclose(g[1], -g[0], rtol=0, atol=0)
        # repulsive: reference uses 0.001 + s where the loss has s
        P = _one_term(1, yi, yj)
        g, _ = P.grad()
        P.m = 0  # remove the (tiny) attractive term of the edge to the far tail
        g0, _ = P.grad()
        g = g - g0
        step = oracle.step_delta(1, yi, yj, A, B, GAMMA, 1.0, np.float64)
        if np.all(np.abs(step) < 4):
            want = step * (0.001 + s) / s
            assert np.allclose(-g[0], want, rtol=1e-10, atol=1e-14)
            assert np.allclose(g[1], -g[0], rtol=1e-12, atol=0)


def test_negative_samples_are_counter_draws():
    P = _problem(seed=2, n=40, deg=3, m=4, coef=False, step=9)
    # recompute the loss by hand from the counter sampler's draws
    tot = 0.0
    for e in range(P.heads.size):
        i, j = P.heads[e], P.tails[e]
        s = max(np.sum((P.emb[i] - P.emb[j]) ** 2), np.finfo(np.float32).eps)
        tot += np.log1p(A * s ** B)
        for p in range(P.m):
            k = oracle.counter_draw(P.seed, P.step, int(P.keys[e]), p, P.n)
            if k == i:
                continue
            s = max(np.sum((P.emb[i] - P.emb[k]) ** 2), np.finfo(np.float32).eps)
            tot += GAMMA * (np.log1p(A * s ** B) - np.log(A) - B * np.log(s))
    assert abs(tot - P.loss()) <= 1e-10 * abs(tot)


def test_clipped_gradient_is_reference_clipped_step():
    """With clip = 4, minus an attractive term's clipped head gradient is the
    reference's clipped step at alpha = 1, also where the clip engages; a
    clip above every term's size changes nothing."""
    rng = np.random.default_rng(8)
    engaged = 0
    a, b = 2.0, 0.3  # steep near 0: the clip engages for close points
    for _ in range(50):
        yi = rng.uniform(-1, 1, size=2)
        yj = yi + rng.normal(scale=rng.choice([1e-3, 0.1, 3.0, 30.0]), size=2)
        P = oracle.LossProblem(np.array([yi, yj]), [0], [1], a, b, GAMMA, m=0)
        g, _ = P.grad(clip=4.0)
        step = oracle.step_delta(0, yi, yj, a, b, GAMMA, 1.0, np.float64)
        assert np.allclose(-g[0], step, rtol=1e-12, atol=1e-15)
        engaged += int(np.any(np.abs(step) == 4.0))
    assert engaged > 0
    P = _problem(seed=6)
    g0, gab0 = P.grad()
    g1, gab1 = P.grad(clip=1e30)
    assert np.array_equal(g0, g1) and np.array_equal(gab0, gab1)
