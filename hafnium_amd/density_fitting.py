"""The fitted tensor on the GPU: ``B = L^-1 (P|uv)`` with the metric ``(P|Q) = L L^T``.

The reference forms ``B = M^-1/2 (P|uv)`` from an eigendecomposition of the
metric with eigenvalues below 1e-10 of the largest dropped
(hafnium/integrals/density_fitting.py:97-107, :181-197). J and K depend on B
only through ``B^T B = (uv|P) M^-1 (Q|wx)``, which any factor ``B = F (P|uv)``
with ``F^T F = M^-1`` reproduces, so the FP64 Cholesky factor (rocSOLVER
dpotrf + rocBLAS dtrsm) gives the same J and K whenever the metric is
positive definite (no eigenvalue would be dropped). A metric that is not
raises ``BackendError``; the eigenvalue-dropping route is then the
reference's.
"""

from __future__ import annotations

import ctypes

import numpy as np

from hafnium_amd.errors import InputError


def _torch():
    import torch

    return torch


def df_tensor(dense, metric, three_center, *, return_factor: bool = False):
    """``L^-1 (P|uv)`` ``[aux, ao, ao]`` as a device tensor, symmetrised in (u, v).

    The triangular solve treats the (u, v) and (v, u) columns separately, so
    its rounding leaves an asymmetry of order 1e-12 relative on larger
    molecules; (B + B^T) / 2 removes it (the packed storage keeps one
    triangle, and its symmetry check is exact). ``return_factor`` also
    returns the Cholesky factor (lower triangle of an [aux, aux] device
    tensor)."""
    from hafnium_amd import _lib

    torch = _torch()
    m = torch.as_tensor(np.asarray(metric) if isinstance(metric, np.ndarray) else metric,
                        dtype=torch.float64).to(dense.device).contiguous()
    t = torch.as_tensor(np.asarray(three_center) if isinstance(three_center, np.ndarray) else three_center,
                        dtype=torch.float64).to(dense.device).contiguous()
    naux = m.shape[0]
    if tuple(m.shape) != (naux, naux):
        raise InputError("the metric is square")
    if t.ndim != 3 or t.shape[0] != naux or t.shape[1] != t.shape[2]:
        raise InputError("three-centre integrals have shape [aux, ao, ao]")
    n = t.shape[1]
    factor = dense.cholesky(m)
    out = dense.lower_solve(factor, t.reshape(naux, n * n)).reshape(naux, n, n)
    _lib.check(dense._lib.hfjk_symmetrize_batched(dense._ctx.handle, n, naux, ctypes.c_void_p(out.data_ptr()),
                                                 dense._stream()))
    return (out, factor) if return_factor else out


def fitted_coulomb_exchange(metric, three_center, *, precision: str = "fp64", device=None):
    """An ``MI355XCoulombExchange`` of the fitted tensor built on the GPU; the metric's Cholesky
    factor stays on it as ``metric_factor`` (used to turn fitted coefficients into charges)."""
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange, _Context
    from hafnium_amd.linalg import Dense

    torch = _torch()
    dev = torch.device(device) if device is not None else torch.device("cuda", 0)
    ctx = _Context(dev.index or 0)
    dense = Dense(ctx, dev)
    b, factor = df_tensor(dense, metric, three_center, return_factor=True)
    out = MI355XCoulombExchange(b, precision=precision, device=dev)
    out.metric_factor = factor
    out.fit_dense = dense
    return out


EXACT_RANK_THRESHOLD = 1e-13
"""Eigenvalues of the [n^2, n^2] integral matrix at or below this fraction of the largest are
treated as its null space (the antisymmetric products and one-centre linear dependencies, which
sit at the rounding level, about 1e-16 relative)."""


def exact_coulomb_exchange(eri, *, precision: str = "fp64", device=None):
    """An ``MI355XCoulombExchange`` that contracts with exact four-centre integrals.

    The drop-in for the reference's ``ExactCoulombExchange(eri)``
    (hafnium/scf/coulomb_exchange.py:74-121: ``J = einsum("uvwx,wx", eri, D)``,
    ``K = einsum("uwxv,wx", eri, D)``). The integral matrix ``V[(uv), (wx)]`` is
    symmetric positive semidefinite, so ``V = sum_P l_P u_P u_P^T`` (rocSOLVER dsyevd on the
    GPU) and ``B[P] = sqrt(l_P) u_P`` over the eigenvalues above ``EXACT_RANK_THRESHOLD`` of the
    largest is an exact factor (``sum_P B[P,uv] B[P,wx] = (uv|wx)``): the DF kernels then give the
    exact J and K, to the threshold. For atoms and small molecules (``n^4`` doubles).
    """
    from hafnium_amd import _lib
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange, _Context
    from hafnium_amd.linalg import Dense

    torch = _torch()
    dev = torch.device(device) if device is not None else torch.device("cuda", 0)
    e = torch.as_tensor(np.asarray(eri) if isinstance(eri, np.ndarray) else eri, dtype=torch.float64)
    n = e.shape[0]
    if e.ndim != 4 or tuple(e.shape) != (n, n, n, n):
        raise InputError("the four-centre integrals have shape [ao, ao, ao, ao]")
    ctx = _Context(dev.index or 0)
    dense = Dense(ctx, dev)
    w, v = dense.eigh(e.to(dev).reshape(n * n, n * n).contiguous())
    if dense.info() != 0:
        raise InputError(f"the integral matrix could not be diagonalised (dsyevd info {dense.info()})")
    top = float(w.abs().max())
    keep = torch.nonzero(w > EXACT_RANK_THRESHOLD * top).flatten()
    if float(w.min()) < -1e-10 * top:
        raise InputError("the four-centre integrals are not a positive semidefinite Coulomb matrix")
    b = (v[:, keep] * w[keep].sqrt()).T.contiguous().reshape(-1, n, n)
    _lib.check(dense._lib.hfjk_symmetrize_batched(dense._ctx.handle, n, b.shape[0], ctypes.c_void_p(b.data_ptr()),
                                                 dense._stream()))
    return MI355XCoulombExchange(b, precision=precision, device=dev)
