"""NumPy restatement of the tiled packed storage of a symmetric ``[aux, ao, ao]`` tensor.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Each ``B[P]`` is cut into ``T x T`` tiles (``T = 32``, zero padding past
``nao``); only tiles ``(r, c)`` with ``r >= c`` are stored, at tile index
``r (r + 1) / 2 + c``, each row-major and complete (the diagonal tiles hold
both triangles). ``P_elems = nt (nt + 1) / 2 * T * T`` with
``nt = ceil(nao / T)``.
"""

from __future__ import annotations

import numpy as np

TILE = 32


def n_tiles(nao: int) -> int:
    return (nao + TILE - 1) // TILE


def packed_elems(nao: int) -> int:
    nt = n_tiles(nao)
    return nt * (nt + 1) // 2 * TILE * TILE


def pack(b: np.ndarray) -> np.ndarray:
    """``[naux, nao, nao]`` symmetric → ``[naux, packed_elems]`` (lower tiles)."""
    naux, n, _ = b.shape
    nt = n_tiles(n)
    npad = nt * TILE
    full = np.zeros((naux, npad, npad), dtype=b.dtype)
    full[:, :n, :n] = b
    out = np.empty((naux, nt * (nt + 1) // 2, TILE, TILE), dtype=b.dtype)
    for r in range(nt):
        for c in range(r + 1):
            out[:, r * (r + 1) // 2 + c] = full[:, r * TILE:(r + 1) * TILE, c * TILE:(c + 1) * TILE]
    return out.reshape(naux, -1)


def unpack(packed: np.ndarray, nao: int) -> np.ndarray:
    """Inverse of :func:`pack`, mirroring the lower tiles into the upper ones."""
    naux = packed.shape[0]
    nt = n_tiles(nao)
    npad = nt * TILE
    tiles = packed.reshape(naux, nt * (nt + 1) // 2, TILE, TILE)
    full = np.zeros((naux, npad, npad), dtype=packed.dtype)
    for r in range(nt):
        for c in range(r + 1):
            t = tiles[:, r * (r + 1) // 2 + c]
            full[:, r * TILE:(r + 1) * TILE, c * TILE:(c + 1) * TILE] = t
            if r != c:
                full[:, c * TILE:(c + 1) * TILE, r * TILE:(r + 1) * TILE] = np.transpose(t, (0, 2, 1))
    return full[:, :nao, :nao]


def weights(density: np.ndarray) -> np.ndarray:
    """Packed weights ``W`` with ``<B[P], D> = <pack(B)[P], W>`` for symmetric B.

    Off-diagonal tiles carry ``D + D^T`` (both triangles of B meet there);
    diagonal tiles carry ``D`` itself.
    """
    n = density.shape[0]
    nt = n_tiles(n)
    npad = nt * TILE
    d = np.zeros((npad, npad))
    d[:n, :n] = density
    sym = d + d.T
    out = np.empty((nt * (nt + 1) // 2, TILE, TILE))
    for r in range(nt):
        for c in range(r + 1):
            src = d if r == c else sym
            out[r * (r + 1) // 2 + c] = src[r * TILE:(r + 1) * TILE, c * TILE:(c + 1) * TILE]
    return out.reshape(-1)
