"""Seeded synthetic DF-J/K inputs, bit-identical on the host and on the GPU.

The workload generator of the benchmark configurations: a fitted tensor
``B[Q, m, n] = B[Q, n, m] = z * exp(-|m - n| / 20)`` with ``z ~ N(0, 1)``
drawn for ``m <= n``, occupied orbitals ``C`` (the Q factor of a reduced QR
of a seeded Gaussian ``[n, o]`` matrix) and the density ``D = C C^T``.

B is too large for host memory at the large configurations (90 GB in FP64
at nao=1500, naux=5000), so every draw is a pure function of its indices:
a Philox4x32-10 block (Salmon, Moraes, Dror and Shaw, "Parallel random
numbers: as easy as 1, 2, 3", SC 2011) keyed by the seed, whose counter is
``(m, n, Q, tag)``, gives two 53-bit uniforms, and the Box-Muller transform
turns them into one normal deviate. The logarithm, the sine and the
exponential are this module's own polynomial evaluations built from IEEE
additions, multiplications, divisions and square roots only, so the NumPy
version here and the HIP kernel ``hfjk_generate`` (compiled with FMA
contraction off) produce the same bits; no libm call is involved.
"""

from __future__ import annotations

import numpy as np

__all__ = [
    "TAG_TENSOR",
    "TAG_ORBITALS",
    "philox4x32",
    "uniform_pair",
    "normal_from_counters",
    "decay",
    "tensor_upper",
    "tensor_full",
    "tensor_slice_full",
    "occupied_orbitals",
    "jk_inputs",
]

TAG_TENSOR = 0x4A4B0001
"""Fourth Philox counter word of the draws of the fitted tensor."""
TAG_ORBITALS = 0x4A4B0002
"""Fourth Philox counter word of the draws of the Gaussian orbital matrix."""

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = 0x9E3779B9
_W1 = 0xBB67AE85
_MASK32 = np.uint64(0xFFFFFFFF)
_SHIFT32 = np.uint64(32)

LN2 = 0.6931471805599453
TWO_PI = 6.283185307179586
DECAY_LENGTH = 20.0


def philox4x32(c0, c1, c2, c3, seed: int):
    """Ten Philox4x32 rounds of the counters ``(c0, c1, c2, c3)`` under ``seed``.

    Inputs are broadcastable integer arrays below 2**32; returns four
    ``uint64`` arrays holding 32-bit words.
    """
    x0 = np.asarray(c0, dtype=np.uint64)
    x1 = np.asarray(c1, dtype=np.uint64)
    x2 = np.asarray(c2, dtype=np.uint64)
    x3 = np.asarray(c3, dtype=np.uint64)
    x0, x1, x2, x3 = np.broadcast_arrays(x0, x1, x2, x3)
    k0 = seed & 0xFFFFFFFF
    k1 = (seed >> 32) & 0xFFFFFFFF
    for _ in range(10):
        p0 = x0 * _M0
        p1 = x2 * _M1
        hi0, lo0 = p0 >> _SHIFT32, p0 & _MASK32
        hi1, lo1 = p1 >> _SHIFT32, p1 & _MASK32
        x0, x1, x2, x3 = (
            hi1 ^ x1 ^ np.uint64(k0),
            lo1,
            hi0 ^ x3 ^ np.uint64(k1),
            lo0,
        )
        k0 = (k0 + _W0) & 0xFFFFFFFF
        k1 = (k1 + _W1) & 0xFFFFFFFF
    return x0, x1, x2, x3


def uniform_pair(c0, c1, c2, c3, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """``u1`` in (0, 1] and ``u2`` in [0, 1), 53 bits each, from one Philox block."""
    r0, r1, r2, r3 = philox4x32(c0, c1, c2, c3, seed)
    scale = 1.0 / 9007199254740992.0  # 2**-53
    a = ((r0 >> np.uint64(5)) << np.uint64(26)) | (r1 >> np.uint64(6))
    b = ((r2 >> np.uint64(5)) << np.uint64(26)) | (r3 >> np.uint64(6))
    u1 = 1.0 - a.astype(np.float64) * scale
    u2 = b.astype(np.float64) * scale
    return u1, u2


def _log(u: np.ndarray) -> np.ndarray:
    """Natural logarithm of ``u > 0`` from ``ln m = 2 atanh((m-1)/(m+1))``."""
    m, e = np.frexp(u)  # u = m 2**e, m in [0.5, 1)
    low = m < 0.7071067811865476
    m = np.where(low, m * 2.0, m)
    e = np.where(low, e - 1, e).astype(np.float64)
    s = (m - 1.0) / (m + 1.0)
    s2 = s * s
    p = 1.0 / 23.0
    for k in (21.0, 19.0, 17.0, 15.0, 13.0, 11.0, 9.0, 7.0, 5.0, 3.0):
        p = p * s2 + 1.0 / k
    p = p * s2 + 1.0
    return e * LN2 + (s + s) * p


def _sin(x: np.ndarray) -> np.ndarray:
    """Sine of ``|x| <= pi/2`` by its Taylor series to ``x**23``."""
    x2 = x * x
    # Coefficients (-1)**((k-1)/2) / k! for k = 23, 21, ..., 3, then 1.
    p = -1.0 / float(_factorial(23))
    for k in (21, 19, 17, 15, 13, 11, 9, 7, 5, 3):
        p = p * x2 + (-1.0 if (k // 2) % 2 else 1.0) / float(_factorial(k))
    p = p * x2 + 1.0
    return x * p


def _factorial(k: int) -> int:
    out = 1
    for i in range(2, k + 1):
        out *= i
    return out


def _cos_two_pi(u: np.ndarray) -> np.ndarray:
    """``cos(2 pi u)`` for ``u`` in [0, 1) through ``-sin(2 pi (v - 1/4))``."""
    v = np.where(u < 0.5, u, 1.0 - u)  # both exact
    w = v - 0.25  # exact for multiples of 2**-53 in [0, 1/2]
    return -_sin(TWO_PI * w)


def _exp_neg(x: np.ndarray) -> np.ndarray:
    """``exp(x)`` for ``x <= 0``: ``2**-k exp(r)`` with ``|r| <= ln2/2``."""
    x = np.asarray(x, dtype=np.float64)
    k = np.floor(-x / LN2 + 0.5)
    r = x + k * LN2
    p = 1.0 / float(_factorial(17))
    for j in range(16, 0, -1):
        p = p * r + 1.0 / float(_factorial(j))
    p = p * r + 1.0
    return np.ldexp(p, (-k).astype(np.int64))


def decay(nao: int) -> np.ndarray:
    """``exp(-d / 20)`` for ``d = 0 .. nao-1``, the envelope of B."""
    d = np.arange(nao, dtype=np.float64)
    return _exp_neg(-(d / DECAY_LENGTH))


def normal_from_counters(c0, c1, c2, c3, seed: int) -> np.ndarray:
    """One Box-Muller deviate per Philox counter."""
    u1, u2 = uniform_pair(c0, c1, c2, c3, seed)
    return np.sqrt(-2.0 * _log(u1)) * _cos_two_pi(u2)


def tensor_upper(seed: int, q: np.ndarray, m: np.ndarray, n: np.ndarray) -> np.ndarray:
    """``B[q, m, n]`` for index arrays with ``m <= n``."""
    z = normal_from_counters(m, n, q, TAG_TENSOR, seed)
    env = decay(int(np.max(n)) + 1 if np.size(n) else 1)
    return z * env[np.asarray(n) - np.asarray(m)]


def tensor_slice_full(seed: int, nao: int, q0: int, q1: int) -> np.ndarray:
    """Full symmetric ``B[q0:q1, :, :]`` as a float64 array."""
    iu, ju = np.triu_indices(nao)
    out = np.empty((q1 - q0, nao, nao), dtype=np.float64)
    env = decay(nao)[ju - iu]
    for row, q in enumerate(range(q0, q1)):
        z = normal_from_counters(iu, ju, q, TAG_TENSOR, seed)
        vals = z * env
        out[row][iu, ju] = vals
        out[row][ju, iu] = vals
    return out


def tensor_full(seed: int, naux: int, nao: int) -> np.ndarray:
    """Full symmetric ``B[naux, nao, nao]``."""
    return tensor_slice_full(seed, nao, 0, naux)


def occupied_orbitals(seed: int, nao: int, nocc: int) -> np.ndarray:
    """Q factor of the reduced QR of a seeded Gaussian ``[nao, nocc]`` matrix."""
    k, i = np.meshgrid(np.arange(nao), np.arange(nocc), indexing="ij")
    gauss = normal_from_counters(k, i, 0xFFFFFFFF, TAG_ORBITALS, seed)
    q, _ = np.linalg.qr(gauss, mode="reduced")
    return np.ascontiguousarray(q)


TAG_BATCH = 0x4A4B0003
"""Fourth Philox counter word of the batch shape draws."""


def batch_shapes(count: int = 512, base_seed: int = 1000, nao_range=(40, 400), aux_ratio: float = 3.5,
                 occ_range=(0.1, 0.3)):
    """``[(seed, nao, naux, nocc)]`` of the batched workload.

    Instance k has seed ``base_seed + k``; ``nao`` is uniform on the integers
    of ``nao_range``, ``naux = round(aux_ratio * nao)`` and
    ``nocc = round(U(occ_range) * nao)`` (at least 1), drawn from one Philox
    block keyed by the instance seed.
    """
    out = []
    lo, hi = nao_range
    for k in range(count):
        seed = base_seed + k
        u1, u2 = uniform_pair(0, 0, 0, TAG_BATCH, seed)
        ua = 1.0 - float(u1)  # in [0, 1)
        nao = lo + min(int(ua * (hi - lo + 1)), hi - lo)
        naux = int(round(aux_ratio * nao))
        frac = occ_range[0] + (occ_range[1] - occ_range[0]) * float(u2)
        nocc = max(1, min(nao, int(round(frac * nao))))
        out.append((seed, nao, naux, nocc))
    return out


def jk_inputs(seed: int, naux: int, nao: int, nocc: int):
    """``(B, C, D)`` of one instance, B in full ``[naux, nao, nao]`` storage."""
    b = tensor_full(seed, naux, nao)
    c = occupied_orbitals(seed, nao, nocc)
    return b, c, c @ c.T
