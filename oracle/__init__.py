"""ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes wrappers over ``oracle/liboracle_umappp.so``, the CPU restatement of
umappp's graph-and-layout hot path (``oracle/umappp_oracle.cpp``).  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module; the product never does.

Parity pins (see tests/test_oracle_*.py):
  neighbor_similarities   reference invariants (neighbor_similarities.cpp:52-228) + root re-solve
  combine_neighbor_sets   brute-force symmetrisation (combine_neighbor_sets.cpp:47-112)
  similarities_to_epochs  closed form, Epochs test (optimize_layout.cpp:58-77)
  Layout (serial)         finite differences of the UMAP losses, closed-form schedule counts,
                          restart equality (optimize_layout.cpp:89-104)
  Layout (parallel)       bit-equality with serial (optimize_layout.cpp:106-125)
  rng kind 0 (mt19937_64 + rejection sampler)  parity unpinned against aarand
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle_umappp.so"

RNG_MT, RNG_COUNTER, RNG_REPLAY = 0, 1, 2

_lib = None
_vp = C.c_void_p
_LOSS_ARGS = [C.c_int64, C.c_int, _vp, C.c_int64, _vp, _vp, _vp, _vp, _vp, C.c_int, C.c_double, C.c_double,
              C.c_double, C.c_uint64, C.c_int]


def build() -> Path:
    subprocess.run(["bash", str(HERE / "build.sh")], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        sig = {
            "orc_neighbor_similarities": (C.c_int, [C.c_int, C.c_int64, _vp, _vp, C.c_double, C.c_double, C.c_double,
                                                    C.c_int, _vp, _vp, C.c_int]),
            "orc_combine_neighbor_sets": (C.c_int64, [C.c_int, C.c_int64, _vp, _vp, _vp, C.c_double, C.c_int64, _vp,
                                                      _vp, _vp]),
            "orc_similarities_to_epochs": (C.c_int64, [C.c_int, C.c_int64, _vp, _vp, _vp, C.c_int, C.c_double, _vp,
                                                       _vp, _vp, _vp, _vp, _vp]),
            "orc_layout_create": (_vp, [C.c_int, C.c_int64, _vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int, C.c_int,
                                        C.c_int, C.c_uint64]),
            "orc_layout_destroy": (None, [_vp]),
            "orc_layout_set_replay": (C.c_int, [_vp, _vp, C.c_int64]),
            "orc_layout_set_record": (C.c_int, [_vp, _vp, C.c_int64]),
            "orc_layout_recorded": (C.c_int64, [_vp]),
            "orc_layout_run": (C.c_int, [_vp, C.c_int, _vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                         C.c_int]),
            "orc_layout_state": (C.c_int, [_vp, _vp, _vp, C.POINTER(C.c_int), _vp, C.c_int64]),
            "orc_layout_set_trace": (C.c_int, [_vp, _vp, C.c_int64]),
            "orc_layout_trace_len": (C.c_int64, [_vp]),
            "orc_step_delta": (C.c_int, [C.c_int, C.c_int, C.c_int, _vp, _vp, C.c_double, C.c_double, C.c_double,
                                         C.c_double, _vp]),
            "orc_epoch_alpha": (C.c_double, [C.c_int, C.c_double, C.c_int, C.c_int]),
            "orc_step_delta_many": (C.c_int, [C.c_int, C.c_int64, _vp, C.c_int, _vp, _vp, C.c_double, C.c_double,
                                              C.c_double, C.c_double, _vp]),
            "orc_counter_draw": (C.c_uint32, [C.c_uint64, C.c_int, C.c_uint64, C.c_uint32, C.c_uint32]),
            "orc_mt19937_64_nth": (C.c_uint64, [C.c_uint64, C.c_int64]),
            "orc_discrete_uniform_stream": (C.c_int, [C.c_uint64, C.c_uint64, C.c_int64, _vp]),
            "orc_gen_normal": (None, [C.c_uint64, C.c_int64, _vp]),
            "orc_gen_uniform01": (None, [C.c_uint64, C.c_int64, _vp]),
            "orc_knn_bruteforce": (None, [C.c_int64, C.c_int, _vp, C.c_int, C.c_int, _vp, _vp]),
            "orc_knn_bruteforce_range": (None, [C.c_int64, C.c_int, _vp, C.c_int, C.c_int64, C.c_int64, C.c_int,
                                                _vp, _vp]),
            "orc_loss": (C.c_double, _LOSS_ARGS),
            "orc_loss_grad": (None, _LOSS_ARGS + [C.c_double, _vp, _vp]),
            "orc_loss_fd": (None, _LOSS_ARGS + [C.c_int64, _vp, _vp, C.c_double, _vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


def _dt(dtype) -> int:
    return 0 if np.dtype(dtype) == np.float32 else 1


def neighbor_similarities(row_ptr, val, local_connectivity=1.0, bandwidth=1.0, min_k_dist_scale=1e-3,
                          use_newton=True, nthreads=1):
    """Returns (weights, rho, sigma); 'val' is not modified."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
    w = np.array(val, copy=True)
    n = rp.size - 1
    rho = np.empty(n)
    sigma = np.empty(n)
    rc = lib().orc_neighbor_similarities(_dt(w.dtype), n, _p(rp), _p(w), local_connectivity, bandwidth,
                                         min_k_dist_scale, 1 if use_newton else 0, _p(rho), _p(sigma), nthreads)
    if rc != 0:
        raise ValueError(f"oracle similarities failed ({rc})")
    return w, rho, sigma


def combine_neighbor_sets(row_ptr, idx, val, mix_ratio=1.0):
    rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
    ix = np.ascontiguousarray(idx, dtype=np.int32)
    v = np.ascontiguousarray(val)
    n = rp.size - 1
    cap = max(1, 2 * int(rp[-1]))
    orp = np.empty(n + 1, dtype=np.int64)
    oidx = np.empty(cap, dtype=np.int32)
    oval = np.empty(cap, dtype=v.dtype)
    m = lib().orc_combine_neighbor_sets(_dt(v.dtype), n, _p(rp), _p(ix), _p(v), float(mix_ratio), cap, _p(orp),
                                        _p(oidx), _p(oval))
    if m < 0:
        raise ValueError(f"oracle combine failed ({m})")
    return orp, oidx[:m].copy(), oval[:m].copy()


def similarities_to_epochs(row_ptr, idx, val, num_epochs, negative_sample_rate):
    """Returns dict with row_ptr, targets, eps, next, next_neg, maxed."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
    ix = np.ascontiguousarray(idx, dtype=np.int32)
    v = np.ascontiguousarray(val)
    n = rp.size - 1
    E = int(rp[-1])
    orp = np.empty(n + 1, dtype=np.int64)
    tg = np.empty(max(E, 1), dtype=np.int32)
    eps = np.empty(max(E, 1), dtype=v.dtype)
    nx = np.empty(max(E, 1), dtype=v.dtype)
    nn = np.empty(max(E, 1), dtype=v.dtype)
    mx = np.zeros(1, dtype=v.dtype)
    m = lib().orc_similarities_to_epochs(_dt(v.dtype), n, _p(rp), _p(ix), _p(v), int(num_epochs),
                                         float(negative_sample_rate), _p(orp), _p(tg), _p(eps), _p(nx), _p(nn), _p(mx))
    if m < 0:
        raise ValueError(f"oracle similarities_to_epochs failed ({m})")
    return dict(row_ptr=orp, targets=tg[:m].copy(), eps=eps[:m].copy(), next=nx[:m].copy(), next_neg=nn[:m].copy(),
                maxed=mx[0])


class Layout:
    """EpochData + sampler state (the reference's Status minus Options)."""

    def __init__(self, epochs: dict, num_epochs: int, negative_sample_rate: float, rng_kind=RNG_MT,
                 seed=1234567890, dtype=None, current_epoch=0):
        dt = np.dtype(dtype) if dtype is not None else epochs["eps"].dtype
        self.dtype = dt
        rp = np.ascontiguousarray(epochs["row_ptr"], dtype=np.int64)
        tg = np.ascontiguousarray(epochs["targets"], dtype=np.int32)
        eps = np.ascontiguousarray(epochs["eps"], dtype=dt)
        nx = np.ascontiguousarray(epochs.get("next", eps), dtype=dt)
        nn = np.ascontiguousarray(epochs.get("next_neg", eps / dt.type(negative_sample_rate)), dtype=dt)
        self.n = rp.size - 1
        self.E = int(rp[-1])
        self.total_epochs = int(num_epochs)
        self._keep = (rp, tg, eps, nx, nn)
        self._h = lib().orc_layout_create(_dt(dt), self.n, _p(rp), _p(tg), _p(eps), _p(nx), _p(nn),
                                          float(negative_sample_rate), int(num_epochs), int(current_epoch),
                                          int(rng_kind), C.c_uint64(int(seed) & (2**64 - 1)))
        self._rec = None

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_layout_destroy(self._h)
            self._h = None

    def set_replay(self, samples):
        s = np.ascontiguousarray(samples, dtype=np.int32)
        self._replay = s
        lib().orc_layout_set_replay(self._h, _p(s), s.size)

    def record(self, cap):
        self._rec = np.zeros(int(cap), dtype=np.int32)
        lib().orc_layout_set_record(self._h, _p(self._rec), self._rec.size)

    def recorded(self):
        m = lib().orc_layout_recorded(self._h)
        return self._rec[: min(m, self._rec.size)].copy()

    def run(self, embedding: np.ndarray, a, b, gamma=1.0, alpha0=1.0, epoch_limit=None, nthreads=1):
        if embedding.dtype != self.dtype or not embedding.flags.c_contiguous:
            raise TypeError("embedding must be C-contiguous of the layout's dtype")
        D = embedding.shape[1] if embedding.ndim == 2 else 1
        lim = self.total_epochs if epoch_limit is None else int(epoch_limit)
        rc = lib().orc_layout_run(self._h, D, _p(embedding), float(a), float(b), float(gamma), float(alpha0), lim,
                                  int(nthreads))
        if rc != 0:
            raise ValueError(f"oracle layout run failed ({rc})")

    def trace(self, cap):
        """Record (kind, head, other) of every step of the next runs."""
        self._trace = np.zeros((int(cap), 3), dtype=np.int64)
        lib().orc_layout_set_trace(self._h, _p(self._trace), int(cap))

    def traced(self):
        m = lib().orc_layout_trace_len(self._h)
        if m > self._trace.shape[0]:
            raise ValueError("oracle trace overflow")
        return self._trace[:m].copy()

    def state(self):
        nx = np.empty(max(self.E, 1), dtype=self.dtype)
        nn = np.empty(max(self.E, 1), dtype=self.dtype)
        cur = C.c_int()
        tal = np.zeros(3 * self.total_epochs, dtype=np.int64)
        lib().orc_layout_state(self._h, _p(nx), _p(nn), C.byref(cur), _p(tal), tal.size)
        return nx[: self.E].copy(), nn[: self.E].copy(), cur.value, tal.reshape(-1, 3)[: cur.value].copy()


def step_delta(kind, head, other, a, b, gamma, alpha, dtype):
    """The head's move for one step (kind 0 attractive, 1 repulsive) from the
    two points' coordinates before it, in the reference's arithmetic."""
    dt = np.dtype(dtype)
    h = np.ascontiguousarray(head, dtype=dt)
    o = np.ascontiguousarray(other, dtype=dt)
    out = np.empty(h.size, dtype=dt)
    lib().orc_step_delta(_dt(dt), int(kind), h.size, _p(h), _p(o), float(a), float(b), float(gamma), float(alpha),
                         _p(out))
    return out


def step_delta_many(kinds, heads, others, a, b, gamma, alpha, dtype):
    """step_delta for many steps at once: kinds (m,), heads and others (m, D)."""
    dt = np.dtype(dtype)
    k = np.ascontiguousarray(kinds, dtype=np.int32)
    h = np.ascontiguousarray(heads, dtype=dt)
    o = np.ascontiguousarray(others, dtype=dt)
    out = np.empty_like(h)
    D = h.shape[1] if h.ndim == 2 else 1
    lib().orc_step_delta_many(_dt(dt), k.size, _p(k), D, _p(h), _p(o), float(a), float(b), float(gamma),
                              float(alpha), _p(out))
    return out


def epoch_alpha(alpha0, n, total, dtype):
    return float(lib().orc_epoch_alpha(_dt(dtype), float(alpha0), int(n), int(total)))


def counter_draw(seed, epoch, edge, p, n):
    return int(lib().orc_counter_draw(C.c_uint64(seed), int(epoch), C.c_uint64(edge), int(p), int(n)))


def mt19937_64_nth(seed, nth):
    return int(lib().orc_mt19937_64_nth(C.c_uint64(seed), int(nth)))


def discrete_uniform_stream(seed, n, count):
    out = np.empty(count, dtype=np.int64)
    lib().orc_discrete_uniform_stream(C.c_uint64(seed), C.c_uint64(n), count, _p(out))
    return out


def gen_normal(seed, count):
    out = np.empty(int(count))
    lib().orc_gen_normal(C.c_uint64(int(seed) & (2**64 - 1)), out.size, _p(out))
    return out


def gen_uniform01(seed, count):
    out = np.empty(int(count))
    lib().orc_gen_uniform01(C.c_uint64(int(seed) & (2**64 - 1)), out.size, _p(out))
    return out


def knn_bruteforce(x: np.ndarray, k: int, nthreads=8):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n, dim = x.shape
    idx = np.empty((n, k), dtype=np.int32)
    dist = np.empty((n, k))
    lib().orc_knn_bruteforce(n, dim, _p(x), k, nthreads, _p(idx), _p(dist))
    return idx, dist


def knn_bruteforce_rows(x: np.ndarray, k: int, q0: int, q1: int, nthreads=8):
    """knn_bruteforce for queries q0 .. q1 - 1 only (against all points):
    (q1 - q0, k) indices and distances, equal to those rows of the full
    search."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    n, dim = x.shape
    if not 0 <= q0 <= q1 <= n:
        raise ValueError("query range outside the data")
    idx = np.empty((q1 - q0, k), dtype=np.int32)
    dist = np.empty((q1 - q0, k))
    lib().orc_knn_bruteforce_range(n, dim, _p(x), k, q0, q1, nthreads, _p(idx), _p(dist))
    return idx, dist


class LossProblem:
    """Inputs of the project's UMAP cross-entropy loss (edge list + negative
    sampling), held as contiguous arrays for the oracle calls."""

    def __init__(self, emb, heads, tails, a, b, gamma=1.0, m=5, seed=0, step=0, keys=None, coef=None, num_neg=None):
        self.emb = np.ascontiguousarray(emb, dtype=np.float64)
        self.n, self.dim = self.emb.shape
        self.heads = np.ascontiguousarray(heads, dtype=np.int32)
        self.tails = np.ascontiguousarray(tails, dtype=np.int32)
        self.keys = None if keys is None else np.ascontiguousarray(keys, dtype=np.int64)
        self.coef = None if coef is None else np.ascontiguousarray(coef, dtype=np.float64)
        self.num_neg = None if num_neg is None else np.ascontiguousarray(num_neg, dtype=np.int32)
        self.a, self.b, self.gamma, self.m, self.seed, self.step = float(a), float(b), float(gamma), int(m), int(seed), int(step)

    def _args(self, emb=None, a=None, b=None):
        e = self.emb if emb is None else np.ascontiguousarray(emb, dtype=np.float64)
        return [self.n, self.dim, _p(e), self.heads.size, _p(self.heads), _p(self.tails), _p(self.keys), _p(self.coef),
                _p(self.num_neg), self.m, self.a if a is None else float(a), self.b if b is None else float(b),
                self.gamma, C.c_uint64(self.seed), self.step]

    def loss(self, emb=None, a=None, b=None) -> float:
        return float(lib().orc_loss(*self._args(emb, a, b)))

    def grad(self, clip=0.0):
        """Closed-form gradient; clip > 0 clips each term's head gradient per coordinate first."""
        g = np.zeros((self.n, self.dim))
        gab = np.zeros(2)
        lib().orc_loss_grad(*self._args(), float(clip), _p(g), _p(gab))
        return g, gab

    def fd(self, points, dims, h=1e-5):
        pts = np.ascontiguousarray(points, dtype=np.int64)
        ds = np.ascontiguousarray(dims, dtype=np.int32)
        out = np.zeros(pts.size)
        lib().orc_loss_fd(*self._args(), pts.size, _p(pts), _p(ds), float(h), _p(out))
        return out


# ---------------------------------------------------------------------------
# Spectral initialisation (dense restatement for small n; numpy)
# ---------------------------------------------------------------------------
def has_multiple_components(row_ptr, idx):
    """spectral_init.hpp:142-168: depth-first walk from observation 0 over the
    symmetric graph; True if it does not reach every observation."""
    rp = np.asarray(row_ptr, dtype=np.int64)
    n = rp.size - 1
    if n == 0:
        return False
    seen = np.zeros(n, dtype=bool)
    seen[0] = True
    stack, count = [0], 1
    while stack:
        c = stack.pop()
        for t in idx[rp[c]:rp[c + 1]]:
            if not seen[t]:
                seen[t] = True
                count += 1
                stack.append(int(t))
    return count != n


def spectral_init_dense(row_ptr, idx, val, num_dim, scale=10.0):
    """spectral_init.hpp:24-140 restated densely: M = 2I - L = I + D^-1/2 W D^-1/2
    (degrees in double, :34-49), the eigenvectors of its num_dim + 1 largest
    eigenvalues without the first (:122-126), scaled so that the largest
    magnitude over all of them is 'scale' (:128-137).  The reference's signs
    come from irlba; here each vector's largest-magnitude entry is made
    positive.  Returns (ok, Y (n, num_dim), eigenvalues of M descending)."""
    rp = np.asarray(row_ptr, dtype=np.int64)
    n = rp.size - 1
    if has_multiple_components(rp, idx) or n <= num_dim + 1:
        return False, None, None
    W = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(rp))
    W[rows, np.asarray(idx)] = np.asarray(val, dtype=np.float64)
    isd = 1.0 / np.sqrt(W.sum(axis=1))
    M = np.eye(n) + isd[:, None] * W * isd[None, :]
    w, V = np.linalg.eigh(M)
    order = np.argsort(-w)
    w, V = w[order], V[:, order]
    Y = V[:, 1:num_dim + 1].copy()
    for c in range(num_dim):
        j = np.argmax(np.abs(Y[:, c]))
        if Y[j, c] < 0:
            Y[:, c] = -Y[:, c]
    Y *= scale / np.max(np.abs(Y))
    return True, Y, w[:num_dim + 1]
