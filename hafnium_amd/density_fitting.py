"""The fitted tensor on the GPU: ``B = F (P|uv)`` with ``F^T F = M^-1`` for the metric ``M = (P|Q)``.

The reference forms ``B = M^-1/2 (P|uv)`` from an eigendecomposition of the
metric with eigenvalues below 1e-10 of the largest dropped
(hafnium/integrals/density_fitting.py:97-107, :181-197, threshold :32). J and
K depend on B only through ``B^T B = (uv|P) M^+ (Q|wx)``, so any factor with
the same ``F^T F`` gives the same J and K:

* "cholesky": ``F = L^-1`` with ``M = L L^T`` (rocSOLVER dpotrf + rocBLAS
  dtrsm), exact when the metric is positive definite and no eigenvalue would
  be dropped;
* "eigh": ``F = diag(w^-1/2) V^T`` over the eigenpairs ``w > 1e-10 w_max``
  (rocSOLVER dsyevd + one GEMM), the reference's ``M^-1/2`` up to the
  orthogonal factor ``V`` and with the dropped rows left out (fewer aux
  functions, the same ``B^T B``);
* "auto" (the default): the metric's eigenvalues decide: "cholesky" when
  none falls below the threshold, "eigh" otherwise.
"""

from __future__ import annotations

import ctypes

import numpy as np

from hafnium_amd.errors import InputError


def _torch():
    import torch

    return torch


DEFAULT_METRIC_THRESHOLD = 1e-10
"""Metric eigenvalues at or below this fraction of the largest are dropped (the reference's
``DEFAULT_METRIC_THRESHOLD``, hafnium/integrals/density_fitting.py:32)."""


def df_tensor(dense, metric, three_center, *, return_factor: bool = False, method: str = "auto",
              threshold: float = DEFAULT_METRIC_THRESHOLD):
    """The fitted tensor ``[aux', ao, ao]`` as a device tensor, symmetrised in (u, v).

    ``method`` is "auto", "cholesky" or "eigh" (module docstring); with "eigh"
    only the retained eigen-directions are rows (``aux' <= aux``). The
    triangular solve and the GEMM treat the (u, v) and (v, u) columns
    separately, so their rounding leaves an asymmetry of order 1e-12 relative
    on larger molecules; (B + B^T) / 2 removes it (the packed storage keeps
    one triangle, and its symmetry check is exact). ``return_factor`` also
    returns ``charges(v)``, the map ``v -> F v`` from ``[aux, k]`` vectors in
    the aux basis to the rows of B (``L^-1 v`` or ``diag(w^-1/2) V^T v``).
    """
    from hafnium_amd import _lib

    torch = _torch()
    if method not in ("auto", "cholesky", "eigh"):
        raise InputError("the fitting method is auto, cholesky or eigh")
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
    flat = t.reshape(naux, n * n)
    keep = None
    if method != "cholesky":
        w, v = dense.eigh(m)
        if dense.info() != 0:
            raise InputError(f"the metric could not be diagonalised (dsyevd info {dense.info()})")
        keep = torch.nonzero(w > threshold * w.max()).flatten()
        if method == "auto" and keep.numel() == naux:
            method, keep = "cholesky", None
        else:
            method = "eigh"
    if method == "cholesky":
        factor = dense.cholesky(m)
        out = dense.lower_solve(factor, flat)
        charges = lambda vec: dense.lower_solve(factor, vec)  # noqa: E731
    else:
        if keep.numel() == 0:
            raise InputError("every metric eigenvalue falls below the threshold")
        rows = (v[:, keep] / w[keep].sqrt()[None, :]).contiguous()  # V diag(w^-1/2): F^T
        out = dense.gemm(rows, flat, ta=True)
        charges = lambda vec: dense.gemm(rows, vec.contiguous(), ta=True)  # noqa: E731
    out = out.reshape(-1, n, n).contiguous()
    _lib.check(dense._lib.hfjk_symmetrize_batched(dense._ctx.handle, n, out.shape[0],
                                                 ctypes.c_void_p(out.data_ptr()), dense._stream()))
    return (out, charges) if return_factor else out


def fitted_coulomb_exchange(metric, three_center, *, precision: str = "fp64", device=None, method: str = "auto",
                            threshold: float = DEFAULT_METRIC_THRESHOLD):
    """An ``MI355XCoulombExchange`` of the fitted tensor built on the GPU (``method`` as in
    :func:`df_tensor`); ``fit_charges(v)`` on it turns ``[aux, k]`` aux-basis vectors ``v = M d`` of
    fitted coefficients ``d`` into the charges of its rows."""
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange, _Context
    from hafnium_amd.linalg import Dense

    torch = _torch()
    dev = torch.device(device) if device is not None else torch.device("cuda", 0)
    ctx = _Context(dev.index or 0)
    dense = Dense(ctx, dev)
    b, charges = df_tensor(dense, metric, three_center, return_factor=True, method=method, threshold=threshold)
    out = MI355XCoulombExchange(b, precision=precision, device=dev)
    out.fit_charges = charges
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
