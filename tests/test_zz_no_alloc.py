"""umx_loss_backward_noalloc against umx_loss_backward, through the C ABI.

The no-allocation variant keeps its working memory in a fixed device array
that every call on the GPU shares, so two calls in a row, the second on a
smaller problem, must each give the allocating entry point's result (the
shared accumulator is cleared by every call), and a problem beyond its limit
is refused with UMX_ERR_CAPACITY.  Gradients are float outputs of double sums
accumulated with atomics in no fixed order, so the two entry points agree to
float rounding (1e-6 relative), not bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A, B, GAMMA = 1.5769285865823632, 0.8950548269037593, 1.0
_vp = C.c_void_p
_ARGS = [C.c_int64, C.c_int, _vp, C.c_int64, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp, C.c_double, C.c_uint64,
         C.c_int, C.c_double, _vp, _vp, _vp, _vp]


@pytest.fixture(scope="module")
def capi():
    import torch

    import umappp_amd

    L = umappp_amd.lib()
    for name in ("umx_loss_backward", "umx_loss_backward_noalloc"):
        fn = getattr(L, name)
        fn.restype = C.c_int
        fn.argtypes = _ARGS
    return L, torch, torch.device("cuda:0")


def _problem(seed, n, dim, deg, m):
    rng = np.random.default_rng(seed)
    emb = rng.uniform(-4, 4, size=(n, dim)).astype(np.float32)
    heads = np.repeat(np.arange(n), deg).astype(np.int32)
    tails = (heads + rng.integers(1, n, size=heads.size)) % n
    coef = rng.uniform(0.1, 1.0, size=heads.size).astype(np.float32)
    nneg = rng.integers(0, m + 2, size=heads.size).astype(np.int32)
    return emb, heads, tails.astype(np.int32), coef, nneg


def _call(capi, name, prob, dim, m, seed, step, clip):
    L, torch, dev = capi
    emb, heads, tails, coef, nneg = prob
    t = [torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in (emb, heads, tails, coef, nneg)]
    ab = torch.tensor([A, B], dtype=torch.float64, device=dev)
    go = torch.tensor([1.5], dtype=torch.float64, device=dev)
    g = torch.full((emb.shape[0], dim), 7.0, dtype=torch.float32, device=dev)  # outputs are overwritten
    gab = torch.full((2,), 7.0, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    rc = getattr(L, name)(emb.shape[0], dim, t[0].data_ptr(), heads.size, t[1].data_ptr(), t[2].data_ptr(), None,
                          t[3].data_ptr(), t[4].data_ptr(), m, ab.data_ptr(), GAMMA, seed, step, clip,
                          go.data_ptr(), g.data_ptr(), gab.data_ptr(), stream)
    return rc, g.cpu().numpy(), gab.cpu().numpy()


@pytest.mark.parametrize("dim,clip", [(2, 0.0), (3, 4.0)])
def test_noalloc_equals_allocating_entry_point(capi, dim, clip):
    for call, (n, deg) in enumerate(((20000, 12), (3000, 7))):  # second call smaller
        prob = _problem(100 + call, n, dim, deg, 5)
        rc0, g0, ab0 = _call(capi, "umx_loss_backward", prob, dim, 5, 77 + call, 3, clip)
        rc1, g1, ab1 = _call(capi, "umx_loss_backward_noalloc", prob, dim, 5, 77 + call, 3, clip)
        assert rc0 == 0 and rc1 == 0
        assert np.all(np.isfinite(g1)) and np.any(g1 != 0)
        np.testing.assert_allclose(g1, g0, rtol=1e-6, atol=1e-6 * float(np.max(np.abs(g0))))
        np.testing.assert_allclose(ab1, ab0, rtol=1e-9)
        print(f"no-alloc backward call {call + 1} (n={n}, D={dim}): max |diff| {np.max(np.abs(g1 - g0)):.2e}")


def test_noalloc_limits_and_errors(capi):
    L, torch, dev = capi
    # beyond the fixed accumulator: refused before any GPU work
    n, dim = (1 << 22) // 2 + 1, 2
    ab = torch.tensor([A, B], dtype=torch.float64, device=dev)
    gab = torch.zeros(2, dtype=torch.float64, device=dev)
    g = torch.zeros(8, dtype=torch.float32, device=dev)
    rc = L.umx_loss_backward_noalloc(n, dim, g.data_ptr(), 0, None, None, None, None, None, 5, ab.data_ptr(), GAMMA,
                                     1, 0, 0.0, None, g.data_ptr(), gab.data_ptr(), None)
    assert rc == -6
    # an edge index outside [0, n) is reported as for the allocating call,
    # and the next call (whose flag starts cleared) succeeds
    prob = list(_problem(5, 500, 2, 4, 3))
    prob[2] = prob[2].copy()
    prob[2][17] = 900
    rc_bad, _, _ = _call(capi, "umx_loss_backward_noalloc", tuple(prob), 2, 3, 1, 0, 0.0)
    assert rc_bad == -2
    good = _problem(6, 500, 2, 4, 3)
    rc0, g0, _ = _call(capi, "umx_loss_backward", good, 2, 3, 1, 0, 0.0)
    rc1, g1, _ = _call(capi, "umx_loss_backward_noalloc", good, 2, 3, 1, 0, 0.0)
    assert rc0 == 0 and rc1 == 0
    np.testing.assert_allclose(g1, g0, rtol=1e-6, atol=1e-6 * float(np.max(np.abs(g0))))
