"""CPU restatement of the learned-guess MLP: TEST INFRASTRUCTURE ONLY.

No counterpart in the reference (its guesses are the core Hamiltonian and the
superposition of atomic densities, hafnium/scf/driver.py:216-246 and
hafnium/scf/atoms.py:76-119); the network is this library's own addition.

Network: ``h_0 = x``, ``z_l = h_{l-1} W_l + b_l``, ``h_l = silu(z_l)`` for the
hidden layers, output ``y = z_L`` (linear). Loss
``L = scale * sum_ij w_j (y_ij - t_ij)^2``.

Two arithmetics:
* ``exact=True``: float64 throughout. Pinned by central finite differences
  of the loss (``finite_difference_check``), so a wrong sign, a dropped
  activation derivative or a transposed product in the backward pass fails.
* ``exact=False``: the kernels' contract, step for step: every GEMM operand
  is rounded to bf16 (nearest even, from its float32 value), products are
  summed in float64 (the kernels sum in float32: the tests allow for that
  difference), and every stored intermediate is rounded to float32.
  Pinned by agreeing with the exact mode to the size of bf16 rounding and
  by ``bf16_round`` reproducing the round-to-nearest-even table of
  hand-picked bit patterns in tests/test_mlp_host.py.
"""

from __future__ import annotations

import numpy as np


def bf16_round(x) -> np.ndarray:
    """Round to bf16 (nearest, ties to even) from the float32 value; returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(layers, x, exact: bool = True):
    """``layers`` = [(W [k, n], b [n]), ...]; returns ``(y, cache)``."""
    h = np.asarray(x, dtype=np.float64)
    cache = []
    last = len(layers) - 1
    for index, (w, b) in enumerate(layers):
        if exact:
            z = h @ w + b
        else:
            z = _f32(bf16_round(h) @ bf16_round(w) + _f32(b))
        cache.append((h, z))
        if index < last:
            h = silu(z) if exact else _f32(silu(z))
        else:
            h = z
    return h, cache


def loss_and_gradients(layers, x, target, col_weights=None, scale: float = 1.0, exact: bool = True):
    """``(loss, [(dW, db), ...])`` of the weighted squared error."""
    y, cache = forward(layers, x, exact)
    w = np.ones(y.shape[1]) if col_weights is None else np.asarray(col_weights, dtype=np.float64)
    d = y - np.asarray(target, dtype=np.float64)
    loss = scale * float(np.sum(w * d * d))
    dz = 2.0 * scale * w * d
    if not exact:
        dz = _f32(dz)
    grads = [None] * len(layers)
    for index in range(len(layers) - 1, -1, -1):
        h_in, z = cache[index]
        wmat = layers[index][0]
        if exact:
            dw = h_in.T @ dz
        else:
            dw = _f32(bf16_round(h_in).T @ bf16_round(dz))
        db = dz.sum(axis=0)
        grads[index] = (dw, db if exact else _f32(db))
        if index > 0:
            dh = dz @ wmat.T if exact else _f32(bf16_round(dz) @ bf16_round(wmat).T)
            z_prev = cache[index - 1][1]
            dz = dh * silu_grad(z_prev) if exact else _f32(dh * silu_grad(z_prev))
    return loss, grads


def finite_difference_check(layers, x, target, col_weights=None, scale: float = 1.0, samples: int = 12,
                            step: float = 1e-6, seed: int = 0) -> float:
    """Largest relative error of the analytic gradient against central differences (exact mode)."""
    rng = np.random.default_rng(seed)
    _, grads = loss_and_gradients(layers, x, target, col_weights, scale, exact=True)
    worst = 0.0
    for _ in range(samples):
        li = int(rng.integers(len(layers)))
        part = int(rng.integers(2))
        arr = layers[li][part]
        idx = tuple(int(rng.integers(s)) for s in arr.shape)
        saved = arr[idx]
        arr[idx] = saved + step
        lp, _ = loss_and_gradients(layers, x, target, col_weights, scale, exact=True)
        arr[idx] = saved - step
        lm, _ = loss_and_gradients(layers, x, target, col_weights, scale, exact=True)
        arr[idx] = saved
        numeric = (lp - lm) / (2.0 * step)
        analytic = grads[li][part][idx]
        worst = max(worst, abs(numeric - analytic) / max(abs(numeric), abs(analytic), 1e-8))
    return worst


def adam_step(params, grads, m1, m2, step: int, lr: float, beta1: float = 0.9, beta2: float = 0.999,
              eps: float = 1e-8, weight_decay: float = 0.0):
    """One Adam step (decoupled weight decay) on flat float64 arrays; returns new (params, m1, m2)."""
    m1 = beta1 * m1 + (1.0 - beta1) * grads
    m2 = beta2 * m2 + (1.0 - beta2) * grads * grads
    lr_t = lr * np.sqrt(1.0 - beta2 ** step) / (1.0 - beta1 ** step)
    params = params - (lr_t * m1 / (np.sqrt(m2) + eps) + lr * weight_decay * params)
    return params, m1, m2
