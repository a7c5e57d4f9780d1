"""The MLP oracle on the CPU: bf16 rounding, finite differences, a planted mistake, Adam."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_ref


def _net(sizes, seed):
    rng = np.random.default_rng(seed)
    return [(rng.standard_normal((k, n)) / np.sqrt(k), 0.1 * rng.standard_normal(n))
            for k, n in zip(sizes[:-1], sizes[1:])]


def test_bf16_round_to_nearest_even():
    ulp = 2.0 ** -7  # bf16 keeps 7 fraction bits
    cases = {
        1.0: 1.0,
        1.0 + ulp / 2: 1.0,                          # tie -> even (mantissa 0)
        1.0 + 1.5 * ulp: 1.0 + 2 * ulp,              # tie -> even (mantissa 2)
        1.0 + ulp / 2 + 2.0 ** -20: 1.0 + ulp,       # above the tie
        1.0 + ulp / 2 - 2.0 ** -20: 1.0,             # below the tie
        -(1.0 + 1.5 * ulp): -(1.0 + 2 * ulp),
        0.0: 0.0,
    }
    for x, want in cases.items():
        assert mlp_ref.bf16_round(np.array([x]))[0] == want, x
    # Every result is a bf16 value: the low 16 bits of its float32 pattern are zero.
    r = mlp_ref.bf16_round(np.random.default_rng(0).standard_normal(1000) * 1e3)
    assert np.all(r.astype(np.float32).view(np.uint32) & 0xFFFF == 0)
    # Relative rounding error is at most half an ulp.
    x = np.random.default_rng(1).uniform(0.5, 1.0, 10000)
    assert np.max(np.abs(mlp_ref.bf16_round(x) - x) / x) <= 2.0 ** -8 + 1e-12


def test_exact_backward_matches_finite_differences():
    layers = _net([12, 16, 16, 5], 3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((9, 12))
    t = rng.standard_normal((9, 5))
    w = rng.uniform(0.5, 2.0, 5)
    assert mlp_ref.finite_difference_check(layers, x, t, w, scale=1.0 / 9, samples=40) < 1e-6


def test_planted_mistake_in_the_backward_is_caught(monkeypatch):
    layers = _net([12, 16, 16, 5], 5)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((9, 12))
    t = rng.standard_normal((9, 5))
    # Drop the z (1 - s) term of silu' (a plausible slip): the check must fail.
    monkeypatch.setattr(mlp_ref, "silu_grad", lambda z: 1.0 / (1.0 + np.exp(-z)))
    assert mlp_ref.finite_difference_check(layers, x, t, samples=40) > 1e-3


def test_rounded_mode_tracks_the_exact_mode():
    layers = _net([64, 48, 48, 7], 7)
    rng = np.random.default_rng(8)
    x = rng.standard_normal((33, 64))
    t = rng.standard_normal((33, 7))
    y_exact, _ = mlp_ref.forward(layers, x, exact=True)
    y_round, _ = mlp_ref.forward(layers, x, exact=False)
    scale = np.abs(y_exact).max()
    err = np.abs(y_round - y_exact).max() / scale
    assert 1e-5 < err < 3e-2  # bf16 operands: visible but small
    _, g_exact = mlp_ref.loss_and_gradients(layers, x, t, exact=True)
    _, g_round = mlp_ref.loss_and_gradients(layers, x, t, exact=False)
    for (dw_e, db_e), (dw_r, db_r) in zip(g_exact, g_round):
        assert np.abs(dw_r - dw_e).max() < 3e-2 * np.abs(dw_e).max()
        assert np.abs(db_r - db_e).max() < 3e-2 * np.abs(db_e).max()


def test_adam_step_closed_form():
    # First step: m1 = (1-b1) g, m2 = (1-b2) g^2, bias-corrected ratio = g / |g| -> p -= lr sign(g).
    g = np.array([0.3, -2.0, 1e-3])
    p, _, _ = mlp_ref.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, lr=0.01, eps=0.0)
    np.testing.assert_allclose(p, -0.01 * np.sign(g), rtol=1e-12)
    # Weight decay is decoupled: zero gradient moves p by -lr wd p only.
    p, _, _ = mlp_ref.adam_step(np.array([2.0]), np.array([0.0]), np.zeros(1), np.zeros(1), 1, lr=0.1,
                                eps=1e-8, weight_decay=0.5)
    assert p[0] == pytest.approx(2.0 - 0.1 * 0.5 * 2.0)
