"""umappp_amd — MI355X (gfx950) implementation of umappp's graph-and-layout hot path.

Python host interface over the C ABI of ``libumappp_amd.so`` (declared in
``include/umappp_amd.h``).  The names, argument meanings and defaults follow
libscran/umappp 3.3.2:

* :func:`neighbor_similarities`   — ``neighbor_similarities.hpp:46-207``
* :func:`combine_neighbor_sets`   — ``combine_neighbor_sets.hpp:15-100``
* :func:`similarities_to_epochs`  — ``optimize_layout.hpp:47-92`` (returns :class:`EpochData`)
* :class:`Options`                 — ``Options.hpp:37-218``
* :func:`initialize`               — ``initialize.hpp:69-119`` (returns :class:`Status`)
* :class:`Status`                  — ``Status.hpp:25-142`` (``run(embedding, epoch_limit)``)
* :func:`find_ab`, :func:`choose_num_epochs` — ``find_ab.hpp:44-168``, ``initialize.hpp:27-47``

The compute runs on the GPU through the library; there is no CPU fallback.
If the library is missing, every call raises :class:`LibraryNotBuilt`.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

__all__ = [
    "LibraryNotBuilt", "UmxError", "NeighborList", "NeighborSimilaritiesOptions", "EpochData", "Options",
    "InitializeMethod", "Status", "neighbor_similarities", "combine_neighbor_sets", "similarities_to_epochs",
    "initialize", "find_ab", "choose_num_epochs", "Graph", "Layout", "library_path", "lib", "find_knn",
]

_HERE = Path(__file__).resolve().parent
_LIBNAME = "libumappp_amd.so"

UMX_F32, UMX_F64 = 0, 1


class LibraryNotBuilt(RuntimeError):
    pass


class UmxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"umx status {code}: {msg}")
        self.code = code


def library_path() -> Path:
    return _HERE / _LIBNAME


_lib_handle = None

_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_vp = C.c_void_p


def _declare(L):
    sig = {
        "umx_last_error": (C.c_char_p, []),
        "umx_version": (C.c_int, []),
        "umx_trim_memory": (C.c_int, []),
        "umx_neighbor_similarities": (C.c_int, [C.c_int, C.c_int64, _vp, _vp, C.c_double, C.c_double, C.c_double,
                                                C.c_int, _vp, _vp]),
        "umx_combine_neighbor_sets": (C.c_int, [C.c_int, C.c_int64, _vp, _vp, _vp, C.c_double, C.c_int64, _vp, _vp,
                                                _vp, _i64p]),
        "umx_graph_create": (C.c_int, [C.c_int, C.c_int64, _vp, _vp, _vp, C.c_int, C.POINTER(_vp)]),
        "umx_graph_neighbor_similarities": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_int]),
        "umx_graph_combine": (C.c_int, [_vp, C.c_double]),
        "umx_graph_info": (C.c_int, [_vp, _i64p, _i64p, C.POINTER(C.c_int)]),
        "umx_graph_export": (C.c_int, [_vp, _vp, _vp, _vp]),
        "umx_graph_export_rho_sigma": (C.c_int, [_vp, _vp, _vp]),
        "umx_graph_destroy": (None, [_vp]),
        "umx_layout_create": (C.c_int, [_vp, C.c_int, C.c_double, C.POINTER(_vp)]),
        "umx_layout_create_csr": (C.c_int, [C.c_int, C.c_int64, _vp, _vp, _vp, C.c_int, C.c_int, C.c_double,
                                            C.POINTER(_vp)]),
        "umx_layout_info": (C.c_int, [_vp, _i64p, _i64p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "umx_layout_export_epochs": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
        "umx_layout_set_params": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_uint64]),
        "umx_layout_set_embedding": (C.c_int, [_vp, _vp, C.c_int]),
        "umx_layout_get_embedding": (C.c_int, [_vp, _vp, C.c_int]),
        "umx_layout_run": (C.c_int, [_vp, C.c_int]),
        "umx_layout_run_serial": (C.c_int, [_vp, C.c_int, _vp, C.c_int64]),
        "umx_layout_tallies": (C.c_int, [_vp, _vp, C.c_int64]),
        "umx_layout_set_mode": (C.c_int, [_vp, C.c_int]),
        "umx_layout_get_mode": (C.c_int, [_vp, C.c_void_p]),
        "umx_layout_launch_shape": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
        "umx_layout_set_phases": (C.c_int, [_vp, C.c_int]),
        "umx_layout_get_phases": (C.c_int, [_vp, _vp, _vp]),
        "umx_layout_set_work_split": (C.c_int, [_vp, C.c_double, C.c_int]),
        "umx_layout_set_tile": (C.c_int, [_vp, C.c_int]),
        "umx_layout_work_items": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int64, _i64p]),
        "umx_layout_set_trace": (C.c_int, [_vp, C.c_int64]),
        "umx_layout_get_trace": (C.c_int, [_vp, _vp, C.c_int64, _i64p]),
        "umx_layout_set_timing": (C.c_int, [_vp, C.c_int]),
        "umx_layout_timing": (C.c_int, [_vp, _f64p, _i64p]),
        "umx_layout_destroy": (None, [_vp]),
        "umx_layout_clone": (C.c_int, [_vp, C.POINTER(C.c_void_p)]),
        "umx_knn_bruteforce": (C.c_int, [C.c_int, C.c_int64, C.c_int, _vp, C.c_int, _vp, _vp, C.c_int]),
        "umx_graph_create_knn": (C.c_int, [C.c_int, C.c_int64, C.c_int, _vp, C.c_int, C.c_int, C.POINTER(_vp)]),
        "umx_graph_components": (C.c_int, [_vp, _i64p]),
        "umx_graph_spectral_init": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int, C.c_double, C.c_uint64, _vp,
                                              C.c_int, C.POINTER(C.c_int), _vp]),
        "umx_graph_spectral_report": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                _f64p, _f64p]),
        "umx_tridiagonal_eigen": (C.c_int, [C.c_int, _vp, _vp, _vp]),
        "umx_graph_create_batch": (C.c_int, [C.c_int, C.c_int32, _vp, _vp, _vp, _vp, C.c_int, C.POINTER(_vp)]),
        "umx_graph_instances": (C.c_int, [_vp, _i32p, _vp]),
        "umx_layout_create_batch": (C.c_int, [_vp, _vp, C.c_double, C.POINTER(_vp)]),
        "umx_layout_instances": (C.c_int, [_vp, _i32p, _vp, _vp]),
        "umx_layout_set_instance_seeds": (C.c_int, [_vp, _vp]),
        "umx_layout_set_placement": (C.c_int, [_vp, C.c_int, C.c_int]),
        "umx_layout_set_graph_block": (C.c_int, [_vp, C.c_int]),
        "umx_layout_autotune": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                          C.POINTER(C.c_int)]),
        "umx_find_ab": (C.c_int, [C.c_double, C.c_double, _f64p, _f64p]),
        "umx_choose_num_epochs": (C.c_int, [C.c_int, C.c_int64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def lib():
    """Loads libumappp_amd.so (raises LibraryNotBuilt when it is absent)."""
    global _lib_handle
    if _lib_handle is None:
        p = library_path()
        if not p.exists():
            raise LibraryNotBuilt(f"{p} is not built; run __graft_entry__.build() or python -m umappp_amd.build")
        # global symbol scope: the libtorch extension of the autograd loss
        # (umappp_amd.autograd) resolves its umx_* calls against this library
        L = C.CDLL(str(p), mode=C.RTLD_GLOBAL)
        _declare(L)
        _lib_handle = L
    return _lib_handle


def _check(rc: int) -> None:
    if rc != 0:
        raise UmxError(rc, lib().umx_last_error().decode())


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _dtype_code(dt) -> int:
    dt = np.dtype(dt)
    if dt == np.float32:
        return UMX_F32
    if dt == np.float64:
        return UMX_F64
    raise TypeError(f"unsupported floating type {dt}")


# ---------------------------------------------------------------------------
# NeighborList (NeighborList.hpp:28-29) as CSR
# ---------------------------------------------------------------------------
class NeighborList:
    """Ragged per-observation lists of (index, value), stored as CSR.

    ``row_ptr`` (int64, n+1), ``idx`` (int32) and ``val`` (float32 or float64).
    Values are distances before :func:`neighbor_similarities` and weights after.
    """

    def __init__(self, row_ptr, idx, val):
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
        self.idx = np.ascontiguousarray(idx, dtype=np.int32)
        v = np.asarray(val)
        if v.dtype not in (np.float32, np.float64):
            v = v.astype(np.float64)
        self.val = np.ascontiguousarray(v)
        if self.row_ptr.ndim != 1 or self.row_ptr.size < 1 or self.row_ptr[0] != 0:
            raise ValueError("row_ptr must be 1-D and start at 0")
        if self.idx.size != self.row_ptr[-1] or self.val.size != self.row_ptr[-1]:
            raise ValueError("idx/val length must equal row_ptr[-1]")

    @classmethod
    def from_knn(cls, indices, distances):
        """From dense (n, k) neighbour index / distance arrays."""
        ind = np.asarray(indices)
        dist = np.asarray(distances)
        n, k = ind.shape
        return cls(np.arange(n + 1, dtype=np.int64) * k, ind.reshape(-1), dist.reshape(-1))

    @classmethod
    def from_lists(cls, rows: Sequence[Sequence[tuple]], dtype=np.float64):
        rp = np.zeros(len(rows) + 1, dtype=np.int64)
        for i, r in enumerate(rows):
            rp[i + 1] = rp[i] + len(r)
        idx = np.fromiter((p[0] for r in rows for p in r), dtype=np.int32, count=int(rp[-1]))
        val = np.fromiter((p[1] for r in rows for p in r), dtype=dtype, count=int(rp[-1]))
        return cls(rp, idx, val)

    def to_lists(self):
        out = []
        for i in range(self.size()):
            s, e = self.row_ptr[i], self.row_ptr[i + 1]
            out.append(list(zip(self.idx[s:e].tolist(), self.val[s:e].tolist())))
        return out

    def size(self) -> int:
        return int(self.row_ptr.size - 1)

    def __len__(self) -> int:
        return self.size()

    def copy(self) -> "NeighborList":
        return NeighborList(self.row_ptr.copy(), self.idx.copy(), self.val.copy())


# ---------------------------------------------------------------------------
# Graph / Layout handles (thin owners of the C handles)
# ---------------------------------------------------------------------------
class Graph:
    """A kNN graph resident on the GPU (umx_graph).

    ``Graph.batch([...])`` builds a batched graph of independent instances
    (one NeighborList each, instance-local indices); the resident graph then
    holds global indices (local + the instance's offset).
    """

    def __init__(self, x: NeighborList, device_ptrs: bool = False, n: int = None, dtype=None,
                 inst_offsets: np.ndarray = None):
        L = lib()
        h = C.c_void_p()
        if inst_offsets is not None:
            off = np.ascontiguousarray(inst_offsets, dtype=np.int64)
            if device_ptrs:
                rp, idx, val = x
                _check(L.umx_graph_create_batch(_dtype_code(dtype), off.size - 1, _ptr(off), C.c_void_p(rp),
                                                C.c_void_p(idx), C.c_void_p(val), 1, C.byref(h)))
            else:
                _check(L.umx_graph_create_batch(_dtype_code(x.val.dtype), off.size - 1, _ptr(off), _ptr(x.row_ptr),
                                                _ptr(x.idx), _ptr(x.val), 0, C.byref(h)))
        elif device_ptrs:
            # x is a tuple of raw device pointers (row_ptr, idx, val)
            rp, idx, val = x
            _check(L.umx_graph_create(_dtype_code(dtype), int(n), C.c_void_p(rp), C.c_void_p(idx),
                                      C.c_void_p(val), 1, C.byref(h)))
        else:
            _check(L.umx_graph_create(_dtype_code(x.val.dtype), x.size(), _ptr(x.row_ptr), _ptr(x.idx),
                                      _ptr(x.val), 0, C.byref(h)))
        self._h = h
        n_, nnz_, dt = C.c_int64(), C.c_int64(), C.c_int()
        _check(L.umx_graph_info(h, C.byref(n_), C.byref(nnz_), C.byref(dt)))
        self.dtype = np.float32 if dt.value == UMX_F32 else np.float64

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib_handle is not None:
            _lib_handle.umx_graph_destroy(h)
            self._h = None

    @classmethod
    def from_data(cls, data: np.ndarray, k: int) -> "Graph":
        """Exact kNN graph of ``data`` (n x dim, float32 or float64) computed on
        the GPU (umx_graph_create_knn); distances in the data's type."""
        x = np.ascontiguousarray(data)
        if x.dtype not in (np.float32, np.float64):
            x = x.astype(np.float64)
        n, dim = x.shape
        g = cls.__new__(cls)
        h = C.c_void_p()
        _check(lib().umx_graph_create_knn(_dtype_code(x.dtype), n, dim, _ptr(x), int(k), 0, C.byref(h)))
        g._h = h
        g.dtype = x.dtype.type
        return g

    @classmethod
    def batch(cls, lists: Sequence[NeighborList]) -> "Graph":
        """Batched graph of independent kNN graphs (all of one float type)."""
        sizes = [x.size() for x in lists]
        off = np.zeros(len(lists) + 1, dtype=np.int64)
        off[1:] = np.cumsum(sizes)
        rp = [np.zeros(1, dtype=np.int64)]
        base = 0
        for x in lists:
            rp.append(x.row_ptr[1:] + base)
            base += int(x.row_ptr[-1])
        cat = NeighborList(np.concatenate(rp), np.concatenate([x.idx for x in lists]),
                           np.concatenate([x.val for x in lists]))
        return cls(cat, inst_offsets=off)

    def instances(self) -> np.ndarray:
        """Observation offsets of the instances (n_inst + 1 entries)."""
        k = C.c_int32()
        _check(lib().umx_graph_instances(self._h, C.byref(k), None))
        off = np.empty(k.value + 1, dtype=np.int64)
        _check(lib().umx_graph_instances(self._h, C.byref(k), _ptr(off)))
        return off

    @property
    def handle(self):
        return self._h

    def info(self):
        n, nnz, dt = C.c_int64(), C.c_int64(), C.c_int()
        _check(lib().umx_graph_info(self._h, C.byref(n), C.byref(nnz), C.byref(dt)))
        return n.value, nnz.value

    def neighbor_similarities(self, local_connectivity=1.0, bandwidth=1.0, min_k_dist_scale=1e-3, use_newton=True):
        _check(lib().umx_graph_neighbor_similarities(self._h, local_connectivity, bandwidth, min_k_dist_scale,
                                                     1 if use_newton else 0))

    def combine(self, mix_ratio=1.0):
        _check(lib().umx_graph_combine(self._h, mix_ratio))

    def export(self) -> NeighborList:
        n, nnz = self.info()
        rp = np.empty(n + 1, dtype=np.int64)
        idx = np.empty(nnz, dtype=np.int32)
        val = np.empty(nnz, dtype=self.dtype)
        _check(lib().umx_graph_export(self._h, _ptr(rp), _ptr(idx), _ptr(val)))
        return NeighborList(rp, idx, val)

    def components(self) -> int:
        """Number of connected components (has_multiple_components is != 1)."""
        c = C.c_int64()
        _check(lib().umx_graph_components(self._h, C.byref(c)))
        return c.value

    def spectral_init(self, num_dim, scale=10.0, jitter=False, jitter_sd=1e-4, seed=1234567890, out=None):
        """spectral_init (spectral_init.hpp:170-207) on the symmetrised graph.

        Returns (success, embedding (n, num_dim) or None, eigenvalues of 2I - L).
        With ``out`` given, it is filled in place on success."""
        n, _ = self.info()
        emb = out if out is not None else np.empty((n, int(num_dim)), dtype=self.dtype)
        ok = C.c_int()
        ev = np.full(int(num_dim) + 1, np.nan)
        _check(lib().umx_graph_spectral_init(self._h, int(num_dim), float(scale), 1 if jitter else 0, float(jitter_sd),
                                             C.c_uint64(int(seed) & (2**64 - 1)), _ptr(emb), 0, C.byref(ok), _ptr(ev)))
        return bool(ok.value), (emb if ok.value else None), ev

    def spectral_report(self):
        """Diagnostics of the last spectral_init: dict with krylov_dim,
        attempts, second_passes (steps re-orthogonalised twice),
        worst_residual and bytes (algorithmic HBM bytes)."""
        kd, at, sp = C.c_int(), C.c_int(), C.c_int()
        wr, by = C.c_double(), C.c_double()
        _check(lib().umx_graph_spectral_report(self._h, C.byref(kd), C.byref(at), C.byref(sp), C.byref(wr),
                                               C.byref(by)))
        return {"krylov_dim": kd.value, "attempts": at.value, "second_passes": sp.value, "worst_residual": wr.value,
                "bytes": by.value}

    def rho_sigma(self):
        n, _ = self.info()
        rho = np.empty(n)
        sigma = np.empty(n)
        _check(lib().umx_graph_export_rho_sigma(self._h, _ptr(rho), _ptr(sigma)))
        return rho, sigma


@dataclass
class EpochData:
    """optimize_layout.hpp:27-45."""
    cumulative_num_edges: np.ndarray
    edge_targets: np.ndarray
    epochs_per_sample: np.ndarray
    epoch_of_next_sample: np.ndarray
    epoch_of_next_negative_sample: np.ndarray
    negative_sample_rate: float
    total_epochs: int
    current_epoch: int = 0


class Layout:
    """EpochData + embedding resident on the GPU (umx_layout)."""

    def __init__(self, graph: Graph = None, num_epochs: int = 500, negative_sample_rate: float = 5.0,
                 csr: NeighborList = None):
        L = lib()
        h = C.c_void_p()
        if graph is not None and np.ndim(num_epochs) == 1:
            ep = np.ascontiguousarray(num_epochs, dtype=np.int32)
            _check(L.umx_layout_create_batch(graph.handle, _ptr(ep), float(negative_sample_rate), C.byref(h)))
            self.dtype = graph.dtype
        elif graph is not None:
            _check(L.umx_layout_create(graph.handle, int(num_epochs), float(negative_sample_rate), C.byref(h)))
            self.dtype = graph.dtype
        else:
            _check(L.umx_layout_create_csr(_dtype_code(csr.val.dtype), csr.size(), _ptr(csr.row_ptr), _ptr(csr.idx),
                                           _ptr(csr.val), 0, int(num_epochs), float(negative_sample_rate),
                                           C.byref(h)))
            self.dtype = csr.val.dtype.type
        self._h = h
        self.negative_sample_rate = float(np.dtype(self.dtype).type(negative_sample_rate))
        self.num_dim = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib_handle is not None:
            _lib_handle.umx_layout_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def clone(self) -> "Layout":
        """Deep copy (umx_layout_clone): schedule, current epoch, tallies,
        parameters, settings and embedding; the copy runs independently of
        this layout and, from the same state, gives the same results bit for
        bit (the reference's Status is copyable, Status.hpp:42-45)."""
        h = C.c_void_p()
        _check(lib().umx_layout_clone(self._h, C.byref(h)))
        other = Layout.__new__(Layout)
        other.__dict__.update({key: val for key, val in self.__dict__.items() if key != "_h"})
        other._h = h
        return other

    def info(self):
        n, nnz, te, ce = C.c_int64(), C.c_int64(), C.c_int(), C.c_int()
        _check(lib().umx_layout_info(self._h, C.byref(n), C.byref(nnz), C.byref(te), C.byref(ce)))
        return n.value, nnz.value, te.value, ce.value

    def epoch_data(self) -> EpochData:
        n, nnz, te, ce = self.info()
        rp = np.empty(n + 1, dtype=np.int64)
        tg = np.empty(nnz, dtype=np.int32)
        eps = np.empty(nnz, dtype=self.dtype)
        nx = np.empty(nnz, dtype=self.dtype)
        nn = np.empty(nnz, dtype=self.dtype)
        _check(lib().umx_layout_export_epochs(self._h, _ptr(rp), _ptr(tg), _ptr(eps), _ptr(nx), _ptr(nn)))
        return EpochData(rp, tg, eps, nx, nn, self.negative_sample_rate, te, ce)

    def set_params(self, num_dim, a, b, gamma=1.0, learning_rate=1.0, seed=1234567890):
        _check(lib().umx_layout_set_params(self._h, int(num_dim), float(a), float(b), float(gamma),
                                           float(learning_rate), C.c_uint64(int(seed) & (2**64 - 1))))
        self.num_dim = int(num_dim)

    def set_embedding(self, emb: np.ndarray = None, device_ptr: int = None):
        if device_ptr is not None:
            _check(lib().umx_layout_set_embedding(self._h, C.c_void_p(device_ptr), 1))
            return
        e = np.ascontiguousarray(emb, dtype=self.dtype)
        _check(lib().umx_layout_set_embedding(self._h, _ptr(e), 0))

    def get_embedding(self, out: np.ndarray = None, device_ptr: int = None):
        if device_ptr is not None:
            _check(lib().umx_layout_get_embedding(self._h, C.c_void_p(device_ptr), 1))
            return None
        n, _, _, _ = self.info()
        if out is None:
            out = np.empty((n, self.num_dim), dtype=self.dtype)
        _check(lib().umx_layout_get_embedding(self._h, _ptr(out), 0))
        return out

    def run(self, epoch_limit: int):
        _check(lib().umx_layout_run(self._h, int(epoch_limit)))

    def run_serial(self, epoch_limit: int, samples: np.ndarray = None):
        if samples is None:
            _check(lib().umx_layout_run_serial(self._h, int(epoch_limit), None, 0))
        else:
            s = np.ascontiguousarray(samples, dtype=np.int32)
            _check(lib().umx_layout_run_serial(self._h, int(epoch_limit), _ptr(s), s.size))

    def tallies(self) -> np.ndarray:
        _, _, te, _ = self.info()
        out = np.zeros(3 * te, dtype=np.int64)
        _check(lib().umx_layout_tallies(self._h, _ptr(out), out.size))
        return out.reshape(te, 3)

    PULL, PUSH = 0, 1

    def set_mode(self, mode: int):
        """Layout.PULL (default: double-buffered, no atomics, bit-reproducible)
        or Layout.PUSH (tails via atomics).  Graphs with one-way entries run in
        push mode only."""
        _check(lib().umx_layout_set_mode(self._h, int(mode)))

    def launch_shape(self):
        """(tile columns, wavefronts per workgroup, placement groups, chunks,
        slots) of the pull kernel's next run."""
        t, w, g = C.c_int(), C.c_int(), C.c_int()
        c, sl = C.c_int64(), C.c_int64()
        _check(lib().umx_layout_launch_shape(self._h, C.byref(t), C.byref(w), C.byref(g), C.byref(c),
                                             C.byref(sl)))
        return t.value, w.value, g.value, c.value, sl.value

    def mode(self) -> int:
        m = C.c_int()
        _check(lib().umx_layout_get_mode(self._h, C.byref(m)))
        return m.value

    def set_work_split(self, fires_per_item: float = 32.0, edges_per_item: int = 32):
        _check(lib().umx_layout_set_work_split(self._h, float(fires_per_item), int(edges_per_item)))
        self._split = (float(fires_per_item), int(edges_per_item))

    def work_split(self):
        """(fires_per_item, edges_per_item) last set on this layout (the
        library's defaults otherwise)."""
        return getattr(self, "_split", (32.0, 32))

    def set_tile(self, tile_cols: int = 0):
        """LDS tile of the pull kernel: 4, 8, 16 or 32 columns, or 0 to choose
        from the number of chunks (the default)."""
        _check(lib().umx_layout_set_tile(self._h, int(tile_cols)))

    def work_items(self, with_phase: bool = False):
        """(begin, length, row, split) arrays of the production kernels' work
        items in run order, padding left out (split: the row has several
        items); with ``with_phase`` also each item's phase and whether it
        starts from the head as an earlier phase of the epoch left it."""
        cnt = C.c_int64()
        _check(lib().umx_layout_work_items(self._h, None, None, None, 0, C.byref(cnt)))
        m = cnt.value
        b = np.empty(m, np.int64)
        ln = np.empty(m, np.int32)
        r = np.empty(m, np.int32)
        _check(lib().umx_layout_work_items(self._h, _ptr(b), _ptr(ln), _ptr(r), m, C.byref(cnt)))
        ru = r.view(np.uint32)
        P, off = self.phases()
        ph = np.repeat(np.arange(P), np.diff(off))
        real = ln > 0
        rows = (ru[real] & 0x7FFFFFFF).astype(np.int64)
        split = np.bincount(rows)[rows] > 1
        out = (b[real], ln[real], rows, split)
        return out + (ph[real], (ru[real] >> 31).astype(bool)) if with_phase else out

    def set_phases(self, phases: int):
        """Phases of a pull-mode epoch: 1, 2, 4 or 8 (default 2)."""
        _check(lib().umx_layout_set_phases(self._h, int(phases)))

    def phases(self):
        """(number of phases, work-item offsets of the phases)."""
        p = C.c_int()
        _check(lib().umx_layout_get_phases(self._h, C.byref(p), None))
        off = np.zeros(p.value + 1, np.int64)
        _check(lib().umx_layout_get_phases(self._h, C.byref(p), _ptr(off)))
        return p.value, off

    def set_trace(self, max_records: int):
        _check(lib().umx_layout_set_trace(self._h, int(max_records)))
        self._trace_cap = int(max_records)

    def get_trace(self, partial: bool = False, with_phase: bool = False):
        """Recorded steps since the last call, as an (m, 3 + 3*num_dim) array:
        kind, head, other, head before, other before, head move.  More steps
        than the capacity is an error unless ``partial``, which returns the
        steps that were recorded (the first ones to claim a record).  With
        ``with_phase`` also each step's launch phase (pull mode; 0 otherwise)."""
        width = 3 + 3 * self.num_dim
        out = np.zeros((self._trace_cap, width))
        total = C.c_int64()
        _check(lib().umx_layout_get_trace(self._h, _ptr(out), self._trace_cap, C.byref(total)))
        if total.value > self._trace_cap and not partial:
            raise UmxError(-6, f"trace overflow: {total.value} steps, capacity {self._trace_cap}")
        out = out[: min(total.value, self._trace_cap)]
        code = out[:, 0].astype(np.int64)
        out[:, 0] = code % 4
        return (out, code // 4) if with_phase else out

    def instances(self):
        """(offsets[n_inst + 1], num_epochs[n_inst]) of the layout's instances."""
        k = C.c_int32()
        _check(lib().umx_layout_instances(self._h, C.byref(k), None, None))
        off = np.empty(k.value + 1, dtype=np.int64)
        ep = np.empty(k.value, dtype=np.int32)
        _check(lib().umx_layout_instances(self._h, C.byref(k), _ptr(off), _ptr(ep)))
        return off, ep

    def set_instance_seeds(self, seeds):
        s = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64))
        _check(lib().umx_layout_set_instance_seeds(self._h, _ptr(s)))

    def set_placement(self, waves: int = 1, groups: int = 0):
        """Wavefronts per workgroup (1, 2, 4) and placement groups (0 = auto, 1 = off)."""
        _check(lib().umx_layout_set_placement(self._h, int(waves), int(groups)))

    def set_graph_block(self, epochs: int):
        """Epochs per captured HIP graph (0 = direct launches)."""
        _check(lib().umx_layout_set_graph_block(self._h, int(epochs)))

    def autotune(self, trial_epochs: int = 3):
        """Runs (or takes from the per-process cache) the launch-shape search; returns (tile, waves, groups)."""
        t, w, g = C.c_int(), C.c_int(), C.c_int()
        _check(lib().umx_layout_autotune(self._h, int(trial_epochs), C.byref(t), C.byref(w), C.byref(g)))
        return t.value, w.value, g.value

    def set_timing(self, on: bool):
        _check(lib().umx_layout_set_timing(self._h, 1 if on else 0))

    def timing(self):
        ms, n = C.c_double(), C.c_int64()
        _check(lib().umx_layout_timing(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value


# ---------------------------------------------------------------------------
# Reference-shaped functions
# ---------------------------------------------------------------------------
@dataclass
class NeighborSimilaritiesOptions:
    """neighbor_similarities.hpp:38-44 (num_threads is accepted and ignored)."""
    local_connectivity: float = 1.0
    bandwidth: float = 1.0
    min_k_dist_scale: float = 1e-3
    num_threads: int = 1


def find_knn(data: np.ndarray, k: int):
    """Exact k nearest neighbours (self excluded, ties to the smaller index) of
    every row of ``data`` on the GPU: (indices int32 (n, k), distances (n, k)
    in the data's float type), each row by increasing distance."""
    x = np.ascontiguousarray(data)
    if x.dtype not in (np.float32, np.float64):
        x = x.astype(np.float64)
    n, dim = x.shape
    idx = np.empty((n, int(k)), dtype=np.int32)
    dist = np.empty((n, int(k)), dtype=x.dtype)
    _check(lib().umx_knn_bruteforce(_dtype_code(x.dtype), n, dim, _ptr(x), int(k), _ptr(idx), _ptr(dist), 0))
    return idx, dist


def neighbor_similarities(x: NeighborList, options: NeighborSimilaritiesOptions = None, use_newton: bool = True,
                          return_rho_sigma: bool = False):
    """In-place distances -> fuzzy weights (neighbor_similarities.hpp:53)."""
    o = options or NeighborSimilaritiesOptions()
    n = x.size()
    rho = np.empty(n) if return_rho_sigma else None
    sigma = np.empty(n) if return_rho_sigma else None
    _check(lib().umx_neighbor_similarities(_dtype_code(x.val.dtype), n, _ptr(x.row_ptr), _ptr(x.val),
                                           o.local_connectivity, o.bandwidth, o.min_k_dist_scale,
                                           1 if use_newton else 0, _ptr(rho), _ptr(sigma)))
    if return_rho_sigma:
        return rho, sigma
    return None


def combine_neighbor_sets(x: NeighborList, mix_ratio: float = 1.0) -> NeighborList:
    """Fuzzy-set symmetrisation (combine_neighbor_sets.hpp:15).  Returns the new
    list and also replaces ``x``'s arrays, mirroring the reference's in-place call."""
    n = x.size()
    cap = max(1, 2 * int(x.row_ptr[-1]))
    orp = np.empty(n + 1, dtype=np.int64)
    oidx = np.empty(cap, dtype=np.int32)
    oval = np.empty(cap, dtype=x.val.dtype)
    nnz = C.c_int64()
    _check(lib().umx_combine_neighbor_sets(_dtype_code(x.val.dtype), n, _ptr(x.row_ptr), _ptr(x.idx), _ptr(x.val),
                                           float(mix_ratio), cap, _ptr(orp), _ptr(oidx), _ptr(oval), C.byref(nnz)))
    m = nnz.value
    x.row_ptr, x.idx, x.val = orp, oidx[:m].copy(), oval[:m].copy()
    return x


def similarities_to_epochs(x: NeighborList, num_epochs: int, negative_sample_rate: float) -> EpochData:
    """optimize_layout.hpp:47."""
    lay = Layout(csr=x, num_epochs=num_epochs, negative_sample_rate=negative_sample_rate)
    return lay.epoch_data()


def trim_memory() -> None:
    """Returns the device buffers the library keeps for reuse on the current
    device to the driver (umx_trim_memory)."""
    _check(lib().umx_trim_memory())


def find_ab(spread: float, min_dist: float):
    """find_ab.hpp:44-168 (host)."""
    a, b = C.c_double(), C.c_double()
    _check(lib().umx_find_ab(float(spread), float(min_dist), C.byref(a), C.byref(b)))
    return a.value, b.value


def choose_num_epochs(num_epochs: Optional[int], size: int) -> int:
    """initialize.hpp:27-47."""
    return int(lib().umx_choose_num_epochs(-1 if num_epochs is None else int(num_epochs), int(size)))


class InitializeMethod:
    SPECTRAL = 0
    RANDOM = 1
    NONE = 2


@dataclass
class Options:
    """Options.hpp:37-218 (defaults unchanged)."""
    local_connectivity: float = 1.0
    bandwidth: float = 1.0
    mix_ratio: float = 1.0
    spread: float = 1.0
    min_dist: float = 0.1
    a: Optional[float] = None
    b: Optional[float] = None
    repulsion_strength: float = 1.0
    initialize_method: int = InitializeMethod.SPECTRAL
    initialize_random_on_spectral_fail: bool = True
    initialize_spectral_scale: float = 10.0
    initialize_spectral_jitter: bool = False
    initialize_spectral_jitter_sd: float = 0.0001
    initialize_random_scale: float = 10.0
    initialize_seed: int = 9876543210
    num_epochs: Optional[int] = None
    learning_rate: float = 1.0
    negative_sample_rate: float = 5.0
    num_neighbors: int = 15
    optimize_seed: int = 1234567890
    num_threads: int = 1
    num_threads_spectral: int = 1
    num_threads_optimize: int = 1


class Status:
    """Status.hpp:25-142.  ``run`` reads the caller's embedding, runs the
    epochs on the GPU and writes the result back, like the reference."""

    def __init__(self, layout: Layout, options: Options, num_dim: int):
        self._layout = layout
        self._options = options
        self._num_dim = int(num_dim)
        self._layout.set_params(num_dim, options.a, options.b, options.repulsion_strength, options.learning_rate,
                                options.optimize_seed)

    def num_dimensions(self) -> int:
        return self._num_dim

    def epoch(self) -> int:
        return self._layout.info()[3]

    def num_epochs(self) -> int:
        return self._layout.info()[2]

    def num_observations(self) -> int:
        return self._layout.info()[0]

    def get_epoch_data(self) -> EpochData:
        return self._layout.epoch_data()

    @property
    def layout(self) -> Layout:
        return self._layout

    def __copy__(self) -> "Status":
        """A deep copy, as the reference's value copy (Status.hpp:42-45): the
        copy owns a clone of the device state and runs independently."""
        other = Status.__new__(Status)
        other._layout = self._layout.clone()
        other._options = self._options
        other._num_dim = self._num_dim
        return other

    def __deepcopy__(self, memo) -> "Status":
        return self.__copy__()

    def run(self, embedding: np.ndarray, epoch_limit: Optional[int] = None) -> None:
        if epoch_limit is None:
            epoch_limit = self.num_epochs()
        if embedding.dtype != self._layout.dtype or not embedding.flags.c_contiguous:
            raise TypeError("embedding must be a C-contiguous array of the graph's floating type")
        self._layout.set_embedding(embedding)
        self._layout.run(epoch_limit)
        self._layout.get_embedding(out=embedding.reshape(-1, self._num_dim))


def initialize(x: NeighborList, num_dim: int, embedding: np.ndarray, options: Options = None) -> Status:
    """initialize.hpp:69-119 on the GPU: similarities, union, initial
    coordinates, (a, b), epoch count and schedule.  SPECTRAL runs the GPU
    spectral initialisation (falling back to RANDOM when it fails and
    ``initialize_random_on_spectral_fail`` is set, initialize.hpp:79-93);
    RANDOM fills uniform(-scale, scale) draws; NONE keeps ``embedding``."""
    import copy

    opt = copy.deepcopy(options) if options is not None else Options()
    g = Graph(x)
    g.neighbor_similarities(opt.local_connectivity, opt.bandwidth, 1e-3, True)
    g.combine(opt.mix_ratio)
    method = opt.initialize_method
    use_random = method == InitializeMethod.RANDOM
    if method == InitializeMethod.SPECTRAL:
        flat = embedding.reshape(-1, num_dim)
        work = np.empty(flat.shape, dtype=g.dtype)
        ok, _, _ = g.spectral_init(num_dim, opt.initialize_spectral_scale, opt.initialize_spectral_jitter,
                                   opt.initialize_spectral_jitter_sd, opt.initialize_seed, out=work)
        if ok:
            flat[:] = work.astype(embedding.dtype)
        use_random = opt.initialize_random_on_spectral_fail and not ok
    if use_random:
        rng = np.random.default_rng(opt.initialize_seed)
        embedding.reshape(-1)[:] = rng.uniform(-opt.initialize_random_scale, opt.initialize_random_scale,
                                               size=embedding.size).astype(embedding.dtype)
    if opt.a is None or opt.b is None:
        opt.a, opt.b = find_ab(opt.spread, opt.min_dist)
    opt.num_epochs = choose_num_epochs(opt.num_epochs, x.size())
    lay = Layout(graph=g, num_epochs=opt.num_epochs, negative_sample_rate=opt.negative_sample_rate)
    return Status(lay, opt, num_dim)
