"""The autograd UMAP loss (HIP forward / backward behind a libtorch
torch::autograd::Function) against the FP64 oracle.

* loss value equal to the oracle's double sum (1e-12 relative);
* embedding gradient equal to the oracle's closed form (float32 rounding:
  1e-6 relative) and, at the differentiable-loss configuration (200,000 x 50
  float32 mixture, fuzzy graph k = 30, out_dim = 3, all edges with weight
  coefficients, 5 negative samples per edge), to FP64 central finite
  differences on 1,000 random coordinates within 1e-4 relative;
* a and b gradients equal to the oracle's closed form (1e-9 relative) and to
  finite differences of the oracle's loss;
* per term, minus the gradient with respect to the head equals the
  reference's update direction (optimize_layout.hpp:145-176): attractive
  exactly, repulsive up to the factor s / (0.001 + s);
* a layout driven through the autograd path keeps the plain layout's
  quality.
"""
import numpy as np
import pytest

import oracle
from _data import knn_lists, mixture

pytestmark = pytest.mark.gpu

A, B, GAMMA = 1.5769285865823632, 0.8950548269037593, 1.0


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    import umappp_amd.autograd as ag

    ag.extension()
    return torch, torch.device("cuda:0")


def _random_problem(seed, n=400, dim=3, deg=8, m=5):
    rng = np.random.default_rng(seed)
    emb = rng.uniform(-4, 4, size=(n, dim)).astype(np.float32)
    heads = np.repeat(np.arange(n), deg).astype(np.int32)
    tails = np.array([rng.choice(np.delete(np.arange(n), i), size=deg, replace=False) for i in range(n)]).reshape(-1)
    coef = rng.uniform(0.1, 1.0, size=heads.size).astype(np.float32)
    nneg = rng.integers(0, m + 2, size=heads.size).astype(np.int32)
    keys = rng.permutation(10 * heads.size)[:heads.size].astype(np.int64)
    return emb, heads, tails.astype(np.int32), coef, nneg, keys


def _to(torch, dev, *arrs):
    return [torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in arrs]


@pytest.mark.parametrize("dim", [1, 2, 3, 5])
def test_loss_and_gradient_match_oracle(torch_dev, dim):
    torch, dev = torch_dev
    from umappp_amd.autograd import umap_loss

    emb, heads, tails, coef, nneg, keys = _random_problem(dim, dim=dim)
    y, h, t, c, nn, k = _to(torch, dev, emb, heads, tails, coef, nneg, keys)
    y.requires_grad_(True)
    a = torch.tensor(A, dtype=torch.float64, device=dev, requires_grad=True)
    b = torch.tensor(B, dtype=torch.float64, device=dev, requires_grad=True)
    loss = umap_loss(y, a, b, h, t, keys=k, coef=c, num_neg=nn, gamma=1.3, seed=77, step=4)
    loss.backward()
    P = oracle.LossProblem(emb.astype(np.float64), heads, tails, A, B, 1.3, seed=77, step=4, keys=keys,
                           coef=coef.astype(np.float64), num_neg=nneg)
    want = P.loss()
    assert abs(float(loss.detach()) - want) <= 1e-12 * abs(want)
    g_ref, gab_ref = P.grad()
    g = y.grad.cpu().numpy().astype(np.float64)
    scale = np.max(np.abs(g_ref))
    assert np.all(np.abs(g - g_ref) <= 1e-6 * np.abs(g_ref) + 1e-7 * scale)
    assert abs(float(a.grad) - gab_ref[0]) <= 1e-9 * abs(gab_ref[0])
    assert abs(float(b.grad) - gab_ref[1]) <= 1e-9 * abs(gab_ref[1])


def test_gradient_with_shuffled_edges(torch_dev):
    """Edges in random order (equal heads in separate runs of a wavefront):
    same gradient as the oracle."""
    torch, dev = torch_dev
    from umappp_amd.autograd import umap_loss

    emb, heads, tails, coef, nneg, keys = _random_problem(31, n=200, deg=12)
    perm = np.random.default_rng(1).permutation(heads.size)
    heads, tails, coef, nneg, keys = heads[perm], tails[perm], coef[perm], nneg[perm], keys[perm]
    heads[::7] = heads[0]  # a heavy head spread across many wavefronts
    ok = heads != tails
    heads, tails, coef, nneg, keys = heads[ok], tails[ok], coef[ok], nneg[ok], keys[ok]
    y, h, t, c, nn, k = _to(torch, dev, emb, heads, tails, coef, nneg, keys)
    y.requires_grad_(True)
    umap_loss(y, A, B, h, t, keys=k, coef=c, num_neg=nn, seed=8).backward()
    P = oracle.LossProblem(emb.astype(np.float64), heads, tails, A, B, 1.0, seed=8, keys=keys,
                           coef=coef.astype(np.float64), num_neg=nneg)
    g_ref, _ = P.grad()
    g = y.grad.cpu().numpy().astype(np.float64)
    scale = np.max(np.abs(g_ref))
    assert np.all(np.abs(g - g_ref) <= 1e-6 * np.abs(g_ref) + 1e-7 * scale)


@pytest.mark.parametrize("seed", [2**63 + 5, 2**64 - 1])
def test_seed_with_top_bit(torch_dev, seed):
    """Seeds of 2^63 and more reach the sampler unchanged (as the layout's
    uint64 optimize_seed does): the loss and gradient equal the oracle's at
    that seed, not at the seed with its top bit cleared."""
    torch, dev = torch_dev
    from umappp_amd.autograd import umap_loss

    emb, heads, tails, coef, nneg, keys = _random_problem(12, n=150, dim=2, deg=6)
    y, h, t, c, nn, k = _to(torch, dev, emb, heads, tails, coef, nneg, keys)
    y.requires_grad_(True)
    loss = umap_loss(y, A, B, h, t, keys=k, coef=c, num_neg=nn, seed=seed, step=2)
    loss.backward()
    P = oracle.LossProblem(emb.astype(np.float64), heads, tails, A, B, 1.0, seed=seed, step=2, keys=keys,
                           coef=coef.astype(np.float64), num_neg=nneg)
    assert abs(float(loss.detach()) - P.loss()) <= 1e-12 * abs(P.loss())
    Q = oracle.LossProblem(emb.astype(np.float64), heads, tails, A, B, 1.0, seed=seed & (2**63 - 1), step=2,
                           keys=keys, coef=coef.astype(np.float64), num_neg=nneg)
    assert Q.loss() != P.loss()
    g_ref, _ = P.grad()
    g = y.grad.cpu().numpy().astype(np.float64)
    assert np.all(np.abs(g - g_ref) <= 1e-6 * np.abs(g_ref) + 1e-7 * np.max(np.abs(g_ref)))


@pytest.mark.parametrize("n,deg", [(401, 7), (333, 5), (70, 3)])
def test_gradient_partial_last_wavefront(torch_dev, n, deg):
    """Edges in row order whose count is not a multiple of 64: the last
    wavefront is partly past the end, and its last run of equal heads ends
    next to those lanes.  Every coordinate of the gradient equals the
    oracle's (a head reduction that read the idle lanes would not)."""
    torch, dev = torch_dev
    from umappp_amd.autograd import umap_loss

    emb, heads, tails, coef, nneg, keys = _random_problem(5 * n + deg, n=n, deg=deg)
    assert heads.size % 64 != 0
    y, h, t, c, nn, k = _to(torch, dev, emb, heads, tails, coef, nneg, keys)
    y.requires_grad_(True)
    umap_loss(y, A, B, h, t, keys=k, coef=c, num_neg=nn, seed=21).backward()
    P = oracle.LossProblem(emb.astype(np.float64), heads, tails, A, B, 1.0, seed=21, keys=keys,
                           coef=coef.astype(np.float64), num_neg=nneg)
    g_ref, _ = P.grad()
    g = y.grad.cpu().numpy().astype(np.float64)
    scale = np.max(np.abs(g_ref))
    assert np.all(np.abs(g - g_ref) <= 1e-6 * np.abs(g_ref) + 1e-7 * scale)


def test_gradient_only_pass(torch_dev):
    """value=False skips the loss kernel: the result is 0 and the gradients
    (embedding, a, b) are the full pass's, also with zero-coefficient edges
    (which the kernels skip) in the list."""
    torch, dev = torch_dev
    from umappp_amd.autograd import umap_loss

    emb, heads, tails, coef, nneg, keys = _random_problem(44, n=300, deg=6)
    coef[::3] = 0
    grads = []
    for value in (True, False):
        y, h, t, c, nn, k = _to(torch, dev, emb, heads, tails, coef, nneg, keys)
        y.requires_grad_(True)
        a = torch.tensor(A, dtype=torch.float64, device=dev, requires_grad=True)
        b = torch.tensor(B, dtype=torch.float64, device=dev, requires_grad=True)
        loss = umap_loss(y, a, b, h, t, keys=k, coef=c, num_neg=nn, seed=3, value=value)
        assert value or float(loss) == 0.0
        loss.backward()
        grads.append((y.grad.cpu().numpy(), float(a.grad), float(b.grad)))
    P = oracle.LossProblem(emb.astype(np.float64), heads, tails, A, B, 1.0, seed=3, keys=keys,
                           coef=coef.astype(np.float64), num_neg=nneg)
    g_ref, gab_ref = P.grad()
    # atomics add in arbitrary order: both passes within float32 rounding of the oracle
    for g, ga, gb in grads:
        scale = np.max(np.abs(g_ref))
        assert np.all(np.abs(g - g_ref) <= 1e-6 * np.abs(g_ref) + 1e-7 * scale)
        assert abs(ga - gab_ref[0]) <= 1e-9 * abs(gab_ref[0]) and abs(gb - gab_ref[1]) <= 1e-9 * abs(gab_ref[1])


def test_clipped_gradient_matches_oracle(torch_dev):
    """Per-term clipping (clip = 4) on a steep curve with close points, where
    the clip engages: equal to the oracle's clipped gradient."""
    torch, dev = torch_dev
    from umappp_amd.autograd import umap_loss

    emb, heads, tails, coef, nneg, keys = _random_problem(21, n=300, dim=2)
    emb = (emb * 0.05).astype(np.float32)  # crowd the points together
    y, h, t, c, nn, k = _to(torch, dev, emb, heads, tails, coef, nneg, keys)
    y.requires_grad_(True)
    umap_loss(y, 2.0, 0.3, h, t, keys=k, coef=c, num_neg=nn, seed=3, clip=4.0).backward()
    P = oracle.LossProblem(emb.astype(np.float64), heads, tails, 2.0, 0.3, 1.0, seed=3, keys=keys,
                           coef=coef.astype(np.float64), num_neg=nneg)
    g_ref, _ = P.grad(clip=4.0)
    g_free, _ = P.grad()
    assert not np.allclose(g_ref, g_free)  # the clip engaged
    g = y.grad.cpu().numpy().astype(np.float64)
    scale = np.max(np.abs(g_ref))
    assert np.all(np.abs(g - g_ref) <= 1e-6 * np.abs(g_ref) + 1e-7 * scale)


def test_ab_gradient_matches_finite_differences(torch_dev):
    torch, dev = torch_dev
    from umappp_amd.autograd import umap_loss

    emb, heads, tails, coef, nneg, keys = _random_problem(9)
    y, h, t, c, nn, k = _to(torch, dev, emb, heads, tails, coef, nneg, keys)
    a = torch.tensor(A, dtype=torch.float64, device=dev, requires_grad=True)
    b = torch.tensor(B, dtype=torch.float64, device=dev, requires_grad=True)
    umap_loss(y, a, b, h, t, keys=k, coef=c, num_neg=nn, seed=5).backward()
    P = oracle.LossProblem(emb.astype(np.float64), heads, tails, A, B, 1.0, seed=5, keys=keys,
                           coef=coef.astype(np.float64), num_neg=nneg)
    hstep = 1e-6
    fa = (P.loss(None, A + hstep, B) - P.loss(None, A - hstep, B)) / (2 * hstep)
    fb = (P.loss(None, A, B + hstep) - P.loss(None, A, B - hstep)) / (2 * hstep)
    assert abs(float(a.grad) - fa) <= 1e-6 * abs(fa)
    assert abs(float(b.grad) - fb) <= 1e-6 * abs(fb)


def test_term_gradient_is_reference_update(torch_dev):
    """Single terms on the GPU: minus the head gradient is the reference's
    step at alpha = 1 (attractive), and the repulsive step times
    (0.001 + s) / s."""
    torch, dev = torch_dev
    from umappp_amd.autograd import umap_loss

    rng = np.random.default_rng(3)
    key = next(kk for kk in range(1000) if oracle.counter_draw(0, 0, kk, 0, 3) == 1)
    for _ in range(10):
        yi = rng.uniform(-2, 2, size=2)
        yj = yi + rng.uniform(-3, 3, size=2)
        s = float(np.sum((yi - yj) ** 2))
        y = torch.tensor(np.array([yi, yj]), dtype=torch.float32, device=dev, requires_grad=True)
        h, t = _to(torch, dev, np.array([0], np.int32), np.array([1], np.int32))
        umap_loss(y, A, B, h, t, m=0).backward()
        yi32, yj32 = y.detach().cpu().numpy().astype(np.float64)
        step = oracle.step_delta(0, yi32, yj32, A, B, 1.0, 1.0, np.float64)
        if np.all(np.abs(step) < 4):
            assert np.allclose(-y.grad.cpu().numpy()[0], step, rtol=1e-6, atol=1e-7)
        # repulsive: edge 0 -> 2 (far away) whose first negative sample is 1
        far = np.array([yi, yj, yi + 1e3])
        y = torch.tensor(far, dtype=torch.float32, device=dev, requires_grad=True)
        h, t, kk = _to(torch, dev, np.array([0], np.int32), np.array([2], np.int32), np.array([key], np.int64))
        umap_loss(y, A, B, h, t, keys=kk, m=1, gamma=GAMMA).backward()
        g1 = y.grad.cpu().numpy()[1].astype(np.float64)  # the sampled point only feels the repulsion
        yv = y.detach().cpu().numpy().astype(np.float64)
        s = float(np.sum((yv[0] - yv[1]) ** 2))
        step = oracle.step_delta(1, yv[0], yv[1], A, B, GAMMA, 1.0, np.float64)
        if np.all(np.abs(step) < 4):
            assert np.allclose(g1, step * (0.001 + s) / s, rtol=1e-6, atol=1e-7)


def test_bad_index_is_an_error(torch_dev):
    torch, dev = torch_dev
    from umappp_amd.autograd import umap_loss

    y = torch.zeros((10, 2), dtype=torch.float32, device=dev)
    h, t = _to(torch, dev, np.array([0, 3], np.int32), np.array([1, 10], np.int32))
    with pytest.raises(RuntimeError, match="outside"):
        umap_loss(y, A, B, h, t)


_C5 = {}


def _c5(torch, dev, umx):
    """Differentiable-loss configuration: 200,000 x 50 float32 mixture
    (30 components), exact kNN k = 30, fuzzy graph (GPU similarities + union),
    out_dim = 3, uniform(-10, 10) embedding."""
    if not _C5:
        import bench

        x, _ = mixture(5, 200_000, 50, 30, dtype=np.float32)
        kidx, kdist = bench.knn_gpu(x, 30, dev)
        n = x.shape[0]
        rp = np.arange(n + 1, dtype=np.int64) * 30
        g = umx.Graph(umx.NeighborList(rp, kidx.cpu().numpy().reshape(-1), kdist.cpu().numpy().reshape(-1)))
        g.neighbor_similarities()
        g.combine(1.0)
        nl = g.export()
        heads = np.repeat(np.arange(n, dtype=np.int32), np.diff(nl.row_ptr))
        emb = np.random.default_rng(55).uniform(-10, 10, size=(n, 3)).astype(np.float32)
        _C5.update(n=n, heads=heads, tails=nl.idx, w=nl.val.astype(np.float32), emb=emb, rp=nl.row_ptr)
    return _C5


def test_c5_gradient_vs_fp64_finite_differences(torch_dev, umx):
    torch, dev = torch_dev
    from umappp_amd.autograd import umap_loss

    c5 = _c5(torch, dev, umx)
    y, h, t, c = _to(torch, dev, c5["emb"], c5["heads"], c5["tails"], c5["w"])
    y.requires_grad_(True)
    a = torch.tensor(A, dtype=torch.float64, device=dev, requires_grad=True)
    b = torch.tensor(B, dtype=torch.float64, device=dev, requires_grad=True)
    loss = umap_loss(y, a, b, h, t, coef=c, m=5, seed=2024, step=0)
    loss.backward()
    P = oracle.LossProblem(c5["emb"].astype(np.float64), c5["heads"], c5["tails"], A, B, 1.0, m=5, seed=2024,
                           step=0, coef=c5["w"].astype(np.float64))
    rng = np.random.default_rng(1000)
    pts = rng.integers(0, c5["n"], size=1000)
    dims = rng.integers(0, 3, size=1000)
    fd = P.fd(pts, dims, h=1e-4)
    got = y.grad.cpu().numpy().astype(np.float64)[pts, dims]
    floor = 1e-3 * np.median(np.abs(fd))
    rel = np.abs(got - fd) / np.maximum(np.abs(fd), floor)
    print(f"C5 ({c5['heads'].size} edges, 5 negatives each): loss {float(loss):.6e}; gradient vs FP64 central "
          f"differences on 1000 coordinates: max relative error {rel.max():.2e}, median {np.median(rel):.2e}")
    assert rel.max() <= 1e-4
    # a, b against the oracle's closed form (its finite differences are pinned on the CPU)
    _, gab = P.grad()
    assert abs(float(a.grad) - gab[0]) <= 1e-9 * abs(gab[0])
    assert abs(float(b.grad) - gab[1]) <= 1e-9 * abs(gab[1])
    assert abs(float(loss) - P.loss()) <= 1e-11 * abs(float(loss))


def test_autograd_layout_quality(torch_dev, umx):
    """10,000-point mixture: 200 epochs driven through the autograd loss keep
    the plain layout's 15-NN preservation and trustworthiness within 0.01."""
    torch, dev = torch_dev
    from _quality import knn_preservation, trustworthiness
    from umappp_amd.autograd import AutogradLayout

    x, _ = mixture(11, 10000, 30, 10, dtype=np.float64)
    rp, idx, dist = knn_lists(x, 15)
    g = umx.Graph(umx.NeighborList(rp, idx, dist.astype(np.float32)))
    g.neighbor_similarities()
    g.combine(1.0)
    a, b = umx.find_ab(1.0, 0.1)
    lay = umx.Layout(graph=g, num_epochs=200, negative_sample_rate=5.0)
    ed = lay.epoch_data()
    emb0 = np.random.default_rng(42).uniform(-10, 10, size=(10000, 2)).astype(np.float32)
    lay.set_params(2, a, b, 1.0, 1.0, 99)
    lay.set_embedding(emb0)
    lay.run(200)
    plain = lay.get_embedding()
    al = AutogradLayout(ed.cumulative_num_edges, ed.edge_targets, ed.epochs_per_sample, 200, 5.0, a, b, seed=99)
    y = torch.from_numpy(emb0.copy()).to(dev)
    al.run(y)
    auto = y.cpu().numpy()
    # the schedule it ran is the plain layout's
    assert np.array_equal(np.asarray(al.fires), lay.tallies()[:, 0])
    hi = idx.reshape(-1, 15)
    qp = knn_preservation(x, plain, 15, high_idx=hi)
    qa = knn_preservation(x, auto, 15, high_idx=hi)
    tp = trustworthiness(x, plain, 15)
    ta = trustworthiness(x, auto, 15)
    print(f"autograd layout vs plain: preservation {qa:.4f} vs {qp:.4f}, trustworthiness {ta:.4f} vs {tp:.4f}")
    assert np.all(np.isfinite(auto))
    assert abs(qa - qp) <= 0.01 and abs(ta - tp) <= 0.01


def test_autograd_layout_at_full_c5_size(torch_dev, umx):
    """Config C5 at full size (200,000 x 50 mixture, k = 30, D = 3): 200 epochs
    driven through the autograd loss (one loss and gradient step per phase,
    as the plain kernel's phases) against the plain layout kernel from the
    same start and schedule: 15-NN preservation within 0.01 of it, and a
    final UMAP cross-entropy (all edges, 5 counter-sampler negatives each)
    no more than 5% above it.  The bound is one-sided: the autograd step is
    the loss's full gradient, which also moves each negative sample away
    from its head (the reference's update moves only the head,
    optimize_layout.hpp:160-176), so it may end lower."""
    torch, dev = torch_dev
    from umappp_amd.autograd import AutogradLayout, edge_heads, umap_loss

    n, k, D, T = 200_000, 30, 3, 200
    x, _ = mixture(5, n, 50, 30)
    idx, dist = umx.find_knn(x, k)
    g = umx.Graph(umx.NeighborList.from_knn(idx, dist))
    g.neighbor_similarities()
    g.combine(1.0)
    a, b = umx.find_ab(1.0, 0.1)
    lay = umx.Layout(graph=g, num_epochs=T, negative_sample_rate=5.0)
    ed = lay.epoch_data()
    emb0 = np.random.default_rng(5).uniform(-10, 10, size=(n, D)).astype(np.float32)
    lay.set_params(D, a, b, 1.0, 1.0, 99)
    lay.set_embedding(emb0)
    lay.run(T)
    plain = torch.from_numpy(lay.get_embedding()).to(dev)
    al = AutogradLayout(ed.cumulative_num_edges, ed.edge_targets, ed.epochs_per_sample, T, 5.0, a, b, seed=99)
    auto = torch.from_numpy(emb0.copy()).to(dev)
    al.run(auto)
    assert np.array_equal(np.asarray(al.fires), lay.tallies()[:, 0])
    assert bool(torch.isfinite(auto).all())

    def preservation(y):
        yc = y.double()
        sq = (yc * yc).sum(1)
        q = torch.arange(0, n, 97, device=dev)
        d2 = sq[q, None] + sq[None, :] - 2.0 * (yc[q] @ yc.T)
        d2[torch.arange(q.numel(), device=dev), q] = float("inf")
        low = torch.topk(d2, 15, largest=False).indices.cpu().numpy()
        hi = idx[q.cpu().numpy(), :15]
        return float(np.mean([len(set(low[r]) & set(hi[r])) for r in range(low.shape[0])])) / 15.0

    nl = g.export()
    heads = edge_heads(torch.from_numpy(nl.row_ptr)).to(dev)
    tails = torch.from_numpy(nl.idx).to(dev)
    w = torch.from_numpy(nl.val.astype(np.float32)).to(dev)
    at = torch.tensor(a, dtype=torch.float64, device=dev)
    bt = torch.tensor(b, dtype=torch.float64, device=dev)

    def cross_entropy(y):
        return float(umap_loss(y, at, bt, heads, tails, coef=w, m=5, seed=4242, step=0)) / float(w.double().sum())

    pp, pa = preservation(plain), preservation(auto)
    cp, ca = cross_entropy(plain), cross_entropy(auto)
    print(f"C5 autograd-driven layout vs plain: 15-NN preservation {pa:.4f} vs {pp:.4f}, "
          f"cross-entropy {ca:.5f} vs {cp:.5f}")
    assert abs(pa - pp) <= 0.01
    assert ca <= 1.05 * cp
