"""A small MLP trained and evaluated by the library's own HIP kernels.

``Mlp(sizes)``: ``z_l = h_{l-1} W_l + b_l``, ``h_l = silu(z_l)`` for hidden
layers, linear output. Parameters live in one flat FP32 buffer (the Adam
master copy) with a BF16 copy that the forward and backward GEMMs read
("BF16 weights, FP32 accumulation"); activations and gradients are FP32 in
HBM and rounded to bf16 as GEMM operands. Every arithmetic step is a call
of the C ABI (``hfjk_mlp_*``); torch only owns the buffers.

No counterpart in the reference (it has no learned components).
"""

from __future__ import annotations

import ctypes
import math

import numpy as np

from hafnium_amd import _lib
from hafnium_amd.coulomb_exchange import _Context, _ptr, _stream
from hafnium_amd.errors import InputError


def _torch():
    import torch

    return torch


class Mlp:
    """Layer sizes ``[n_in, hidden..., n_out]`` on one GPU."""

    def __init__(self, sizes, *, seed: int = 0, device=None, context=None) -> None:
        torch = _torch()
        if len(sizes) < 2 or any(int(s) < 1 for s in sizes):
            raise InputError("an MLP needs at least an input and an output size, all positive")
        self.sizes = [int(s) for s in sizes]
        self.device = torch.device(device if device is not None else "cuda:0")
        if self.device.type != "cuda":
            raise InputError("the MLP runs on a GPU")
        self._ctx = context if context is not None else _Context(self.device.index or 0)
        self.lib = _lib.load()
        self.offsets = []
        at = 0
        for k, n in zip(self.sizes[:-1], self.sizes[1:]):
            self.offsets.append((at, at + k * n))  # W [k, n], then b [n]
            at += k * n + n
        self.n_params = at
        self.params = torch.zeros(at, dtype=torch.float32, device=self.device)
        self.params16 = torch.zeros(at, dtype=torch.bfloat16, device=self.device)
        self.grads = torch.zeros(at, dtype=torch.float32, device=self.device)
        self.moment1 = torch.zeros(at, dtype=torch.float32, device=self.device)
        self.moment2 = torch.zeros(at, dtype=torch.float32, device=self.device)
        self.step_count = 0
        self.set_params(self.initial_params(seed))
        self._loss = torch.zeros(1, dtype=torch.float32, device=self.device)

    # ------------------------------------------------------------ parameters
    def initial_params(self, seed: int) -> np.ndarray:
        """Gaussian weights with variance 1/fan-in (the output layer scaled by 0.1), zero biases."""
        rng = np.random.default_rng(seed)
        flat = np.zeros(self.n_params)
        last = len(self.sizes) - 2
        for index, ((k, n), (w0, b0)) in enumerate(zip(zip(self.sizes[:-1], self.sizes[1:]), self.offsets)):
            gain = 0.1 if index == last else 1.0
            flat[w0:b0] = rng.standard_normal(k * n) * gain / math.sqrt(k)
        return flat

    def set_params(self, flat) -> None:
        torch = _torch()
        flat = np.asarray(flat, dtype=np.float32)
        if flat.shape != (self.n_params,):
            raise InputError(f"expected {self.n_params} parameters, got {flat.shape}")
        self.params.copy_(torch.from_numpy(flat).to(self.device))
        self._check(self.lib.hfjk_mlp_to_bf16(self._ctx.handle, self.n_params, _ptr(self.params), _ptr(self.params16),
                                              self._stream()))
        self.moment1.zero_()
        self.moment2.zero_()
        self.step_count = 0

    def layers(self) -> list[tuple[np.ndarray, np.ndarray]]:
        """Host copies ``[(W [k, n], b [n]), ...]`` of the FP32 master parameters."""
        flat = self.params.cpu().numpy().astype(np.float64)
        out = []
        for (k, n), (w0, b0) in zip(zip(self.sizes[:-1], self.sizes[1:]), self.offsets):
            out.append((flat[w0:b0].reshape(k, n).copy(), flat[b0:b0 + n].copy()))
        return out

    def gradients(self) -> list[tuple[np.ndarray, np.ndarray]]:
        flat = self.grads.cpu().numpy().astype(np.float64)
        out = []
        for (k, n), (w0, b0) in zip(zip(self.sizes[:-1], self.sizes[1:]), self.offsets):
            out.append((flat[w0:b0].reshape(k, n).copy(), flat[b0:b0 + n].copy()))
        return out

    # ------------------------------------------------------------ helpers
    def _stream(self):
        return _stream(_torch(), self.device)

    def _check(self, status: int) -> None:
        _lib.check(status)

    def _addr(self, tensor, offset_elems: int = 0) -> ctypes.c_void_p:
        return ctypes.c_void_p(tensor.data_ptr() + offset_elems * tensor.element_size())

    def _gemm(self, m, n, k, a, a_off, a_bf16, lda, ta, rows, b, b_off, b_bf16, ldb, tb, bias, c, h=None):
        self._check(self.lib.hfjk_mlp_gemm(
            self._ctx.handle, m, n, k, self._addr(a, a_off), int(a_bf16), lda, int(ta),
            None if rows is None else self._addr(rows[0], rows[1]), self._addr(b, b_off), int(b_bf16), ldb, int(tb),
            None if bias is None else self._addr(bias[0], bias[1]), _ptr(c), n, 0,
            None if h is None else _ptr(h), self._stream()))

    def _rows(self, x, rows):
        torch = _torch()
        if x.dtype != torch.float32 or x.dim() != 2 or x.shape[1] != self.sizes[0] or not x.is_contiguous():
            raise InputError(f"inputs are a contiguous float32 [m, {self.sizes[0]}] tensor")
        if x.device != self.device:
            raise InputError("inputs live on the network's GPU")
        if rows is None:
            return x.shape[0], None
        rows_t, lo, hi = rows
        if rows_t.dtype != torch.int32 or rows_t.device != self.device:
            raise InputError("row indices are an int32 tensor on the network's GPU")
        if not 0 <= lo <= hi <= rows_t.numel():
            raise InputError("row range out of bounds")
        return hi - lo, (rows_t, lo)

    # ------------------------------------------------------------ passes
    def forward(self, x, rows=None, *, keep: bool = False):
        """Outputs ``[m, n_out]`` for inputs ``x [N, n_in]``; ``rows = (int32 tensor, lo, hi)``
        selects the minibatch ``x[rows[lo:hi]]`` (indices checked by the caller to address x)."""
        torch = _torch()
        m, gather = self._rows(x, rows)
        acts = []
        h = x
        for index, ((k, n), (w0, b0)) in enumerate(zip(zip(self.sizes[:-1], self.sizes[1:]), self.offsets)):
            z = torch.empty((m, n), dtype=torch.float32, device=self.device)
            last = index == len(self.offsets) - 1
            h_next = None if last else torch.empty((m, n), dtype=torch.float32, device=self.device)
            self._gemm(m, n, k, h, 0, False, k, False, gather if index == 0 else None, self.params16, w0, True, n,
                       False, (self.params, b0), z, h_next)
            acts.append((h, z))
            h = z if last else h_next
        if keep:
            self._acts = (acts, gather, m)
        return h

    def backward(self, grad_out) -> None:
        """Parameter gradients (into ``self.grads``) from ``d loss / d outputs`` of the last kept forward."""
        torch = _torch()
        acts, gather, m = self._acts
        dz = grad_out
        for index in range(len(self.offsets) - 1, -1, -1):
            k, n = self.sizes[index], self.sizes[index + 1]
            w0, b0 = self.offsets[index]
            h_in, _ = acts[index]
            dw = self.grads[w0:b0]
            # dW = h_in^T dz  (reduction over the minibatch rows)
            self._gemm(k, n, m, h_in, 0, False, k, True, gather if index == 0 else None, dz, 0, False, n, False,
                       None, dw)
            self._check(self.lib.hfjk_mlp_column_sum(self._ctx.handle, m, n, _ptr(dz), n,
                                                     self._addr(self.grads, b0), self._stream()))
            if index > 0:
                dh = torch.empty((m, k), dtype=torch.float32, device=self.device)
                # dh = dz W^T, with W read transposed from its bf16 copy
                self._gemm(m, k, n, dz, 0, False, n, False, None, self.params16, w0, True, n, True, None, dh)
                z_prev = acts[index - 1][1]
                dz_prev = torch.empty_like(dh)
                self._check(self.lib.hfjk_mlp_silu_backward(self._ctx.handle, m * k, _ptr(dh), _ptr(z_prev),
                                                            _ptr(dz_prev), self._stream()))
                dz = dz_prev

    def loss(self, y, target, rows=None, col_weights=None, scale: float = 1.0, grad=None):
        """Device scalar ``scale * sum w_j (y - target[rows])^2``; fills ``grad`` when given."""
        m, n = y.shape
        gather = None
        if rows is not None:
            gather = self._addr(rows[0], rows[1])
        self._check(self.lib.hfjk_mlp_mse(self._ctx.handle, m, n, _ptr(y), n, _ptr(target), target.shape[1], gather,
                                          None if col_weights is None else _ptr(col_weights), float(scale),
                                          None if grad is None else _ptr(grad), _ptr(self._loss), self._stream()))
        return self._loss

    def adam(self, lr: float, *, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8,
             weight_decay: float = 0.0) -> None:
        self.step_count += 1
        t = self.step_count
        lr_t = lr * math.sqrt(1.0 - beta2 ** t) / (1.0 - beta1 ** t)
        self._check(self.lib.hfjk_mlp_adam(self._ctx.handle, self.n_params, _ptr(self.params), _ptr(self.grads),
                                           _ptr(self.moment1), _ptr(self.moment2), _ptr(self.params16), float(lr),
                                           float(lr_t), float(beta1), float(beta2), float(eps), float(weight_decay),
                                           self._stream()))

    def train_step(self, x, target, rows, *, lr: float, col_weights=None, weight_decay: float = 0.0):
        """One minibatch step (forward, mean weighted squared error, backward, Adam); returns the device loss."""
        torch = _torch()
        m, _ = self._rows(x, rows)
        y = self.forward(x, rows, keep=True)
        grad = torch.empty_like(y)
        loss = self.loss(y, target, rows, col_weights, 1.0 / max(m, 1), grad)
        self.backward(grad)
        self.adam(lr, weight_decay=weight_decay)
        return loss
