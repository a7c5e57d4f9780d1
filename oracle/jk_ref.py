"""NumPy restatement of the density-fitted Coulomb and exchange contractions.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows hafnium/scf/coulomb_exchange.py:176-190 (``coulomb``: charges
``c_P = sum_wx B[P,w,x] D[w,x]`` then ``J = sum_P c_P B[P]``; ``exchange``:
``K = sum_P B[P] D B[P]`` as ``(B @ D)`` permuted to ``[ao, aux*ao]`` times
``B`` reshaped to ``[aux*ao, ao]``), and the occupied-orbital factorisation
the reference names as a later optimisation at coulomb_exchange.py:129-133:
``X_P = B[P] C``, ``K = sum_P X_P X_P^T``.
"""

from __future__ import annotations

import numpy as np


def coulomb(b: np.ndarray, density: np.ndarray) -> np.ndarray:
    """``J`` of hafnium/scf/coulomb_exchange.py:176-182 (two GEMVs over B)."""
    naux, n, _ = b.shape
    flat = b.reshape(naux, n * n)
    charges = flat @ density.reshape(n * n)
    return (charges @ flat).reshape(n, n)


def charges(b: np.ndarray, density: np.ndarray) -> np.ndarray:
    """``gamma_P = <B[P], D>``, the first GEMV of :func:`coulomb`."""
    naux, n, _ = b.shape
    return b.reshape(naux, n * n) @ density.reshape(n * n)


def exchange(b: np.ndarray, density: np.ndarray) -> np.ndarray:
    """``K`` of hafnium/scf/coulomb_exchange.py:184-190 (density form)."""
    naux, n, _ = b.shape
    half = b @ density
    left = np.transpose(half, (1, 0, 2)).reshape(n, naux * n)
    return left @ b.reshape(naux * n, n)


def exchange_occupied(b: np.ndarray, orbitals: np.ndarray) -> np.ndarray:
    """``K = sum_P (B[P] C)(B[P] C)^T``, equal to :func:`exchange` for ``D = C C^T``."""
    naux, n, _ = b.shape
    x = b @ orbitals  # [naux, n, o]
    left = np.transpose(x, (1, 0, 2)).reshape(n, -1)
    return left @ left.T


def half_transform(b: np.ndarray, orbitals: np.ndarray) -> np.ndarray:
    """``X[P] = B[P] C`` ``[naux, n, o]``."""
    return b @ orbitals


def jk_streamed(seed_slices, density: np.ndarray, orbitals: np.ndarray | None = None):
    """Accumulate J and K over Q slices of B (J and K are sums over Q).

    ``seed_slices`` yields ``[q, n, n]`` slices; K uses the occupied form
    when ``orbitals`` is given, the density form otherwise.
    """
    n = density.shape[0]
    j = np.zeros((n, n))
    k = np.zeros((n, n))
    for part in seed_slices:
        j += coulomb(part, density)
        k += exchange(part, density) if orbitals is None else exchange_occupied(part, orbitals)
    return j, k


def energies(density: np.ndarray, j: np.ndarray, k: np.ndarray) -> tuple[float, float]:
    """``E_J = tr(D J)/2`` and ``E_K = tr(D K)/2`` on the bare density."""
    return 0.5 * float(np.sum(density * j.T)), 0.5 * float(np.sum(density * k.T))


def max_relative_error(value: np.ndarray, reference: np.ndarray) -> float:
    """Normwise max relative error ``||a - r||_max / ||r||_max``."""
    scale = float(np.max(np.abs(reference)))
    return float(np.max(np.abs(value - reference))) / (scale if scale > 0 else 1.0)


def looped_coulomb_exchange(eri: np.ndarray, density: np.ndarray):
    """The four-index definitions ``J[u,v] = sum (uv|wx) D[w,x]`` and
    ``K[u,v] = sum (uw|xv) D[w,x]`` as explicit loops (hafnium/scf/
    coulomb_exchange.py:53-71; tests/unit/test_coulomb_exchange.py:29-42).

    For tiny ``n`` only: it is the brute-force check of the contractions.
    """
    n = eri.shape[0]
    dtype = np.result_type(eri, density)
    j = np.zeros((n, n), dtype=dtype)
    k = np.zeros((n, n), dtype=dtype)
    for u in range(n):
        for v in range(n):
            sj = 0.0
            sk = 0.0
            for w in range(n):
                for x in range(n):
                    sj += eri[u, v, w, x] * density[w, x]
                    sk += eri[u, w, x, v] * density[w, x]
            j[u, v] = sj
            k[u, v] = sk
    return j, k
