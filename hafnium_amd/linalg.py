"""FP64 dense linear algebra of the SCF on the GPU, through the C ABI.

Thin typed wrappers over ``hfjk_dgemm`` (rocBLAS), ``hfjk_dsyevd``,
``hfjk_dpotrf`` (rocSOLVER), ``hfjk_dtrsm_lower`` and the library's own
elementwise and fixed-order reduction kernels. Tensors are torch FP64 CUDA
tensors used as buffers only: every arithmetic operation here runs in
``libhfjk.so``.
"""

from __future__ import annotations

import ctypes

from hafnium_amd import _lib
from hafnium_amd.errors import BackendError, InputError


def _torch():
    import torch

    return torch


def _p(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class Dense:
    """FP64 dense operations on one device, sharing a context with a J/K builder."""

    def __init__(self, ctx, device) -> None:
        torch = _torch()
        self._ctx = ctx
        self._lib = ctx.lib
        self.device = device
        self._partial = torch.empty(int(self._lib.hfjk_dot_workspace()), dtype=torch.float64, device=device)
        self._info = torch.zeros(1, dtype=torch.int32, device=device)
        # Count of failed diagonalisations since the last reset, kept on the device so that
        # checking it costs no synchronisation of its own (the SCF reads it with its energies).
        self._eig_failures = torch.zeros(1, dtype=torch.float64, device=device)

    def _stream(self):
        return ctypes.c_void_p(_torch().cuda.current_stream(self.device).cuda_stream)

    def _check(self, *tensors):
        for t in tensors:
            if t is None:
                continue
            if t.dtype != _torch().float64 or not t.is_contiguous() or t.device != self.device:
                raise InputError("dense FP64 routines take contiguous float64 tensors on the context's device")

    def gemm(self, a, b, *, ta: bool = False, tb: bool = False, alpha: float = 1.0, beta: float = 0.0, out=None):
        """``out = alpha op(a) op(b) + beta out``."""
        torch = _torch()
        self._check(a, b, out)
        m = a.shape[1] if ta else a.shape[0]
        k = a.shape[0] if ta else a.shape[1]
        kb = b.shape[1] if tb else b.shape[0]
        n = b.shape[0] if tb else b.shape[1]
        if k != kb:
            raise InputError(f"gemm inner dimensions {k} and {kb} differ")
        if out is None:
            out = torch.zeros((m, n), dtype=torch.float64, device=self.device) if beta != 0.0 else torch.empty(
                (m, n), dtype=torch.float64, device=self.device)
        elif tuple(out.shape) != (m, n):
            raise InputError("gemm output has the wrong shape")
        _lib.check(self._lib.hfjk_dgemm(self._ctx.handle, int(ta), int(tb), m, n, k, alpha, _p(a), a.shape[1], _p(b),
                                        b.shape[1], beta, _p(out), n, self._stream()))
        return out

    def eigh(self, a):
        """``(w, v)`` of a symmetric ``a``: ascending ``w``, eigenvectors as the COLUMNS of ``v``."""
        torch = _torch()
        self._check(a)
        n = a.shape[0]
        work = a.clone()
        w = torch.empty(n, dtype=torch.float64, device=self.device)
        e = torch.empty(n, dtype=torch.float64, device=self.device)
        _lib.check(self._lib.hfjk_dsyevd(self._ctx.handle, n, _p(work), _p(w), _p(e), _p(self._info),
                                         self._stream()))
        self._eig_failures += (self._info != 0).to(torch.float64)
        rows = work  # row j = eigenvector j
        return w, self.transpose(rows)

    def info(self) -> int:
        return int(self._info.item())

    @property
    def eig_failures(self):
        """Device float64 ``[1]``: diagonalisations of :meth:`eigh` whose dsyevd info was nonzero
        since the last :meth:`reset_eig_failures`."""
        return self._eig_failures

    def reset_eig_failures(self) -> None:
        self._eig_failures.zero_()

    def transpose(self, a):
        """``a^T`` through a GEMM with the identity-free transposed copy."""
        torch = _torch()
        m, n = a.shape
        eye = torch.eye(m, dtype=torch.float64, device=self.device)
        return self.gemm(a, eye, ta=True)

    def cholesky_solve(self, metric, rhs):
        """``L^-1 rhs`` with ``metric = L L^T`` (``rhs`` ``[m, k]``, overwritten copy returned)."""
        return self.lower_solve(self.cholesky(metric), rhs)

    def cholesky(self, metric):
        """``L`` (lower triangle of the returned copy; the strict upper triangle keeps the input)."""
        self._check(metric)
        m = metric.shape[0]
        fac = metric.clone()
        _lib.check(self._lib.hfjk_dpotrf(self._ctx.handle, m, _p(fac), _p(self._info), self._stream()))
        if self.info() != 0:
            raise BackendError(f"the metric is not positive definite (dpotrf info {self.info()})")
        return fac

    def lower_solve(self, factor, rhs):
        """``L^-1 rhs`` for the lower triangle ``L`` of ``factor`` (``rhs`` ``[m, k]``; a solved copy)."""
        self._check(factor, rhs)
        out = rhs.clone()
        _lib.check(self._lib.hfjk_dtrsm_lower(self._ctx.handle, factor.shape[0], rhs.shape[1], _p(factor), _p(out),
                                              self._stream()))
        return out

    def combine(self, a: float, x, b: float = 0.0, y=None, c: float = 0.0, z=None, out=None):
        """``out = a x + b y + c z``."""
        torch = _torch()
        self._check(x, y, z, out)
        if out is None:
            out = torch.empty_like(x)
        _lib.check(self._lib.hfjk_combine(self._ctx.handle, x.numel(), a, _p(x), b, _p(y), c, _p(z), _p(out),
                                          self._stream()))
        return out

    def antisymmetrize(self, m, out=None):
        """``m - m^T``."""
        torch = _torch()
        self._check(m, out)
        if out is None:
            out = torch.empty_like(m)
        _lib.check(self._lib.hfjk_antisymmetrize(self._ctx.handle, m.shape[0], _p(m), _p(out), self._stream()))
        return out

    def dot(self, x, y, out):
        """``out[0] = sum x*y`` (device scalar slot ``out``, a 1-element view)."""
        self._check(x, y)
        _lib.check(self._lib.hfjk_ddot(self._ctx.handle, x.numel(), _p(x), _p(y), _p(self._partial), _p(out),
                                       self._stream()))
        return out
