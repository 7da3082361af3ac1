"""UMAP cross-entropy loss as a torch autograd function, and a layout driven by it.

:func:`umap_loss` is a ``torch::autograd::Function`` (libtorch C++ extension
``umappp_amd_torch``) whose forward and backward are HIP kernels of
libumappp_amd (``umx_loss_forward`` / ``umx_loss_backward``).  Over an edge
list ``heads[e] -> tails[e]`` of an embedding ``y`` (float32, n x dim)::

    L = sum_e c_e [ log(1 + a s^b)
                    + gamma * sum_{p < m_e, k != head} (log(1 + a s_k^b) - log a - b log s_k) ]

with ``s = max(FLT_EPSILON, |y_head - y_other|^2)``.  Gradients reach ``y``
and the curve parameters ``a`` and ``b``.  Negative sample ``p`` of edge ``e``
is the layout's counter-sampler draw for ``(seed, step, keys[e], p)``; the
backward pass recomputes it.  There is no reference counterpart: the
reference applies per-edge SGD steps (``optimize_layout.hpp:145-176``), and
minus this loss's per-term gradient with respect to the head is exactly the
reference's attractive step direction, and its repulsive one with ``s`` in
place of ``0.001 + s``.

:class:`AutogradLayout` runs the reference's epoch schedule
(``optimize_layout.hpp:116-185``: which edges fire in an epoch and how many
negative samples each draws) but moves the embedding by the gradient of the
epoch's loss, computed through :func:`umap_loss`.
"""
from __future__ import annotations

import importlib.util
from pathlib import Path

import numpy as np

from . import LibraryNotBuilt, lib

_PKG = Path(__file__).resolve().parent
_EXT_NAME = "umappp_amd_torch"
_EXT_PATH = _PKG / "_torch_build" / f"{_EXT_NAME}.so"
_ext = None


def extension():
    """Loads the libtorch extension (raises LibraryNotBuilt when absent)."""
    global _ext
    if _ext is None:
        import torch  # noqa: F401  (the extension needs libtorch loaded)

        lib()  # global scope: the extension calls the library's C ABI
        if not _EXT_PATH.exists():
            raise LibraryNotBuilt(f"{_EXT_PATH} is not built; run __graft_entry__.build()")
        spec = importlib.util.spec_from_file_location(_EXT_NAME, str(_EXT_PATH))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        _ext = mod
    return _ext


def _scalar(x, dev):
    import torch

    if isinstance(x, torch.Tensor):
        return x
    return torch.tensor(float(x), dtype=torch.float64, device=dev)


def umap_loss(emb, a, b, heads, tails, *, keys=None, coef=None, num_neg=None, m=5, gamma=1.0, seed=0, step=0,
              clip=0.0, value=True):
    """UMAP cross-entropy of ``emb`` over the edge list (see the module docstring).

    ``emb``: float32 (n, dim) GPU tensor; ``a``, ``b``: floats or float64
    scalar tensors (gradients flow to tensors that require them);
    ``heads``/``tails``: int32 GPU tensors; ``keys`` (int64, sampler key per
    edge; default the edge's position), ``coef`` (float32) and ``num_neg``
    (int32, negatives per edge; default ``m``) are optional.  ``clip > 0``
    makes the backward pass clip every term's gradient with respect to its
    head to ``[-clip, clip]`` per coordinate before summing (per-sample
    clipping; 4 is the reference's clip); ``clip = 0`` gives the exact
    gradient.  Returns a float64 scalar tensor.  ``value=False`` skips the
    loss kernel and returns 0 (the gradient is unchanged), for callers that
    only differentiate.
    """
    import torch

    dev = emb.device
    empty = lambda dt: torch.empty(0, dtype=dt, device=dev)  # noqa: E731
    stream = torch.cuda.current_stream(dev).cuda_stream
    return extension().umap_loss(
        emb, _scalar(a, dev), _scalar(b, dev), heads, tails,
        keys if keys is not None else empty(torch.int64),
        coef if coef is not None else empty(torch.float32),
        num_neg if num_neg is not None else empty(torch.int32),
        int(m), float(gamma), _as_int64(seed), int(step), float(clip), int(stream), bool(value))


def _as_int64(seed) -> int:
    """A uint64 seed as the int64 with the same bits (two's complement), which
    the extension casts back to uint64 unchanged."""
    s = int(seed) & (2**64 - 1)
    return s - 2**64 if s >= 2**63 else s


def edge_heads(row_ptr):
    """Head of every entry of a CSR (int32), from its row offsets."""
    import torch

    rp = torch.as_tensor(row_ptr)
    counts = (rp[1:] - rp[:-1]).to(torch.int64)
    return torch.repeat_interleave(torch.arange(counts.numel(), device=rp.device, dtype=torch.int32), counts)


class AutogradLayout:
    """Layout epochs whose moves come from the autograd loss.

    The schedule is the reference's (``optimize_layout.hpp``): edge ``j`` fires
    in epoch ``n`` when ``epoch_of_next_sample[j] <= n`` (:141-143) and then
    draws ``num_neg = (int)((n - epoch_of_next_negative_sample[j]) /
    (epochs_per_sample[j] / negative_sample_rate))`` negative samples
    (:160-161), after which both schedule values advance (:178-179).  The
    epoch's loss is :func:`umap_loss` over the firing edges with those sample
    counts (keyed by the edge's schedule index and the epoch, so the samples
    are the plain layout kernel's), and the embedding moves by
    ``-alpha_n * grad`` with the reference's learning-rate decay (:131-133),
    the gradient's terms clipped one by one at the reference's clip value
    (:105-110) in the backward pass, so that a step is the reference's epoch
    of clipped per-edge steps.

    The epoch runs in ``phases`` steps (default 4), as the plain layout's
    epochs run in phases: the firing edges are dealt to phases by their head
    (blocks of 16 observations, block b to phase b mod phases, the plain
    kernel's deal), and each phase's loss is differentiated at the positions
    the earlier phases' steps left, so later phases see fresh positions of
    the points the earlier ones moved (with one phase every term of an epoch
    reads the epoch's starting positions, and heavily connected points
    overshoot: 15-NN preservation 0.0061 against the plain layout's 0.0117
    on the 200k-point k = 30 graph; four phases: 0.0167).

    Unlike the reference, whose negative samples move only the head, the
    loss's repulsive terms also move the sampled point (they are terms of one
    loss); at the reference's gamma this gives the plain layout's quality
    (15-NN preservation and trustworthiness within 0.002 on a 10,000-point
    mixture).
    """

    def __init__(self, row_ptr, targets, epochs_per_sample, num_epochs, negative_sample_rate=5.0, a=None, b=None,
                 gamma=1.0, learning_rate=1.0, seed=1234567890, device="cuda", phases=4):
        import torch

        self.dev = torch.device(device)
        self.rp = torch.as_tensor(np.asarray(row_ptr, dtype=np.int64), device=self.dev)
        self.tails = torch.as_tensor(np.asarray(targets, dtype=np.int32), device=self.dev)
        self.heads = edge_heads(self.rp).contiguous()
        self.eps = torch.as_tensor(np.asarray(epochs_per_sample, dtype=np.float32), device=self.dev)
        self.nsr = float(np.float32(negative_sample_rate))
        self.next = self.eps.clone()
        self.next_neg = self.eps / np.float32(self.nsr)
        self.keys = torch.arange(self.eps.numel(), dtype=torch.int64, device=self.dev)
        self.total = int(num_epochs)
        self.current = 0
        self.a = torch.tensor(float(a), dtype=torch.float64, device=self.dev)
        self.b = torch.tensor(float(b), dtype=torch.float64, device=self.dev)
        self.gamma = float(gamma)
        self.alpha0 = float(learning_rate)
        self.seed = int(seed)
        if int(phases) not in (1, 2, 4, 8):
            raise ValueError("phases must be 1, 2, 4 or 8")
        self.phases = int(phases)
        # phase of every entry's head: blocks of 16 observations dealt in turn
        self.head_phase = ((self.heads >> 4) & (self.phases - 1)).to(torch.int64)
        self._fire_counts = []

    def _alpha(self, n):
        # (float)(alpha0 * (1.0 - epoch / num_epochs)): the quotient in float,
        # the rest in double (optimize_layout.hpp:131-133)
        x = float(np.float32(n) / np.float32(self.total))
        return float(np.float32(float(np.float32(self.alpha0)) * (1.0 - x)))

    def run(self, emb, epoch_limit=None):
        """Runs epochs [current, epoch_limit) on ``emb`` (float32 (n, dim) GPU
        tensor, updated in place).  Each epoch compacts its firing edges
        before the loss (passing the whole edge list with a 0/1 mask instead
        leaves most lanes of every wavefront idle, since about 1 edge in 5
        fires, and was slower) and orders them by phase (a stable sort, so
        each phase keeps the edges' schedule order)."""
        import torch

        limit = self.total if epoch_limit is None else int(epoch_limit)
        epns_all = self.eps / np.float32(self.nsr)
        # epoch numbers as float32 on the device (a fresh host scalar per
        # epoch would be a synchronous copy)
        epochs = torch.arange(limit, dtype=torch.float32, device=self.dev)
        for n in range(self.current, limit):
            ep = epochs[n]
            fire = self.next <= ep
            idx = torch.nonzero(fire).squeeze(1)
            self._fire_counts.append(int(idx.numel()))
            if idx.numel():
                epns = epns_all[idx]
                nneg = torch.trunc((ep - self.next_neg[idx]) / epns).to(torch.int32)
                if self.phases > 1:
                    ph = self.head_phase[idx]
                    order = torch.sort(ph, stable=True).indices
                    sizes = torch.bincount(ph, minlength=self.phases).tolist()
                    pidx, pneg = idx[order], nneg[order]
                else:
                    sizes, pidx, pneg = [int(idx.numel())], idx, nneg
                lo = 0
                for size in sizes:
                    if size:
                        sel = pidx[lo:lo + size]
                        y = emb.detach().requires_grad_(True)
                        loss = umap_loss(y, self.a, self.b, self.heads[sel].contiguous(), self.tails[sel].contiguous(),
                                         keys=self.keys[sel].contiguous(), num_neg=pneg[lo:lo + size].contiguous(),
                                         gamma=self.gamma, seed=self.seed, step=n, clip=4.0, value=False)
                        (g,) = torch.autograd.grad(loss, [y])
                        with torch.no_grad():
                            emb.add_(g, alpha=-self._alpha(n))
                    lo += size
                self.next[idx] += self.eps[idx]
                self.next_neg[idx] += nneg.to(torch.float32) * epns
            self.current = n + 1
        return emb

    @property
    def fires(self):
        """Firing edges (edge-updates) of every epoch run so far."""
        return list(self._fire_counts)
