"""NumPy restatement of the k-point density-fitted Coulomb and exchange contractions.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows hafnium/scf/coulomb_exchange.py:193-394:

* ``KPointDensityFittedCoulombExchange.coulomb`` (:279-288): with the
  zero-transfer blocks ``B[k, k]`` and their signs ``s[k, k]``, the charges
  ``c_P = sum_k' w_k' sum_xw conj(B[k', k', P, x, w]) D(k')[x, w]`` and
  ``J(k) = sum_P s[k, k, P] c_P B[k, k, P]``;
* ``KPointDensityFittedCoulombExchange.exchange`` (:290-297):
  ``K(k) = sum_k' w_k' sum_P s[k, k', P] B[k, k', P] D(k') B[k, k', P]^dagger
  + xi S(k) D(k) S(k)``;
* ``KPointDensityFittedCoulomb.coulomb`` (:375-383): the same J from a
  zero-transfer tensor ``[kpt, aux, ao, ao]`` with one sign per aux function.

``supercell_tensor`` builds, from a real translation-invariant tensor of a
ring of cells, both the Gamma-point tensor of the whole ring and the
k-point tensor of its Bloch transform, so the k-point formulas can be checked
against the molecular ones on the supercell.
"""

from __future__ import annotations

import numpy as np


def coulomb(df_tensor: np.ndarray, signature: np.ndarray, weights: np.ndarray, density: np.ndarray) -> np.ndarray:
    """``J(k)`` ``[kpt, ao, ao]`` (coulomb_exchange.py:279-288)."""
    nk = df_tensor.shape[0]
    diagonal = np.stack([df_tensor[k, k] for k in range(nk)])  # [kpt, aux, ao, ao]
    signs = np.stack([signature[k, k] for k in range(nk)])  # [kpt, aux]
    inner = np.sum(np.conj(diagonal) * density[:, None, :, :], axis=(-2, -1))  # [kpt, aux]
    charges = np.sum(inner * weights[:, None], axis=0)  # [aux]
    signed = signs.astype(np.complex128) * charges[None, :]
    return np.sum(diagonal * signed[:, :, None, None], axis=1)


def exchange(df_tensor, signature, weights, overlap, madelung: float, density: np.ndarray) -> np.ndarray:
    """``K(k)`` ``[kpt, ao, ao]`` of one spin block (coulomb_exchange.py:290-297)."""
    half = df_tensor @ density[None, :, None, :, :]
    signed = half * signature.astype(np.complex128)[..., None, None]
    product = np.sum(signed @ np.conj(np.swapaxes(df_tensor, -1, -2)), axis=2)  # [k1, k2, ao, ao]
    weighted = np.sum(product * weights[None, :, None, None], axis=1)
    return weighted + madelung * (overlap @ density @ overlap)


def coulomb_zero_transfer(df_tensor: np.ndarray, signature: np.ndarray, weights: np.ndarray,
                          density: np.ndarray) -> np.ndarray:
    """``J(k)`` of ``KPointDensityFittedCoulomb`` (coulomb_exchange.py:375-383)."""
    inner = np.sum(np.conj(df_tensor) * density[:, None, :, :], axis=(-2, -1))
    charges = np.sum(inner * weights[:, None], axis=0)
    signed = signature.astype(np.complex128) * charges
    return np.sum(df_tensor * signed[None, :, None, None], axis=1)


def supercell_tensor(beta: np.ndarray):
    """Gamma-point and k-point forms of one translation-invariant fitted tensor of a ring of cells.

    ``beta[P, mu, nu, a, b]`` (real, ``beta[P, mu, nu, a, b] = beta[P, nu, mu, b, a]``) is the fitted
    tensor element of aux function P in cell T, orbital mu in cell T + a and orbital nu in cell
    T + b, cells counted modulo N. Returns ``(b_super, b_k, kpoints)``:

    * ``b_super[(P, T), (mu, Tm), (nu, Tn)] = beta[P, mu, nu, Tm - T, Tn - T]``, the real
      ``[N naux, N nao, N nao]`` tensor of the supercell, index ``(x, T) -> T n_x + x``;
    * ``b_k[k1, k2, P, mu, nu] = sum_{u, a} exp(-i k1 u + i k2 a) beta[P, mu, nu, u, a]`` with
      ``k = 2 pi j / N``, the Bloch blocks for which the k-point formulas with weights 1/N
      reproduce the supercell's Gamma-point J and K under ``D(k)[x, w] = sum_T D_super[(x, 0),
      (w, T)] exp(i k T)`` and ``X(k) = sum_T X_super[(mu, 0), (nu, T)] exp(i k T)``.
    """
    naux, n, _, ncell, _ = beta.shape
    ks = 2.0 * np.pi * np.arange(ncell) / ncell
    cells = np.arange(ncell)
    b_super = np.zeros((ncell * naux, ncell * n, ncell * n))
    for t in range(ncell):
        for tm in range(ncell):
            for tn in range(ncell):
                b_super[t * naux:(t + 1) * naux, tm * n:(tm + 1) * n, tn * n:(tn + 1) * n] = \
                    beta[:, :, :, (tm - t) % ncell, (tn - t) % ncell]
    phase1 = np.exp(-1j * np.outer(ks, cells))  # [k1, u]
    phase2 = np.exp(1j * np.outer(ks, cells))  # [k2, a]
    b_k = np.einsum("iu,ja,pmnua->ijpmn", phase1, phase2, beta)
    return b_super, b_k, ks


def to_supercell(blocks: np.ndarray) -> np.ndarray:
    """``X_super[(mu, Tm), (nu, Tn)] = (1 / N) sum_k exp(i k (Tm - Tn)) X(k)[mu, nu]`` for
    ``blocks[k]`` on the ring's k set."""
    ncell, n, _ = blocks.shape
    ks = 2.0 * np.pi * np.arange(ncell) / ncell
    out = np.zeros((ncell * n, ncell * n), dtype=np.complex128)
    for tm in range(ncell):
        for tn in range(ncell):
            out[tm * n:(tm + 1) * n, tn * n:(tn + 1) * n] = np.tensordot(
                np.exp(1j * ks * (tm - tn)), blocks, axes=(0, 0)) / ncell
    return out


def from_supercell(matrix: np.ndarray, ncell: int) -> np.ndarray:
    """``X(k)[mu, nu] = sum_T X_super[(mu, 0), (nu, T)] exp(i k T)``."""
    n = matrix.shape[0] // ncell
    ks = 2.0 * np.pi * np.arange(ncell) / ncell
    row = matrix[:n, :].reshape(n, ncell, n)  # [mu, T, nu]
    return np.einsum("kt,mtn->kmn", np.exp(1j * np.outer(ks, np.arange(ncell))), row)
