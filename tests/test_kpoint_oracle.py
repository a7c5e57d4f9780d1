"""The k-point oracle pinned by the supercell equivalence: the Bloch blocks of a translation-invariant
ring of cells with weights 1/N give the Gamma-point J and K of the whole ring, which the molecular
oracle (itself pinned by the stored fixtures) computes directly."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import jk_ref, kpoint_ref


def ring(seed, naux=4, nao=3, ncell=3):
    """A random translation-invariant fitted tensor and density of a ring of ``ncell`` cells."""
    rng = np.random.default_rng(seed)
    beta = rng.standard_normal((naux, nao, nao, ncell, ncell))
    beta = 0.5 * (beta + np.transpose(beta, (0, 2, 1, 4, 3)))  # B[P] symmetric in its orbital pairs
    d = rng.standard_normal((nao, nao, ncell))
    d = 0.5 * (d + np.transpose(d, (1, 0, 2))[:, :, (-np.arange(ncell)) % ncell])
    d_super = np.zeros((ncell * nao, ncell * nao))
    for tx in range(ncell):
        for tw in range(ncell):
            d_super[tx * nao:(tx + 1) * nao, tw * nao:(tw + 1) * nao] = d[:, :, (tw - tx) % ncell]
    return beta, d_super


@pytest.mark.parametrize("seed,ncell", [(1, 1), (2, 3), (3, 4)])
def test_kpoint_contractions_reproduce_the_supercell(seed, ncell):
    beta, d_super = ring(seed, ncell=ncell)
    assert np.allclose(d_super, d_super.T)
    b_super, b_k, _ = kpoint_ref.supercell_tensor(beta)
    naux, n = beta.shape[0], beta.shape[1]
    assert np.allclose(b_super, np.transpose(b_super, (0, 2, 1)))
    d_k = kpoint_ref.from_supercell(d_super, ncell)
    np.testing.assert_allclose(kpoint_ref.to_supercell(d_k), d_super, atol=1e-12)
    sig = np.ones((ncell, ncell, naux))
    w = np.full(ncell, 1.0 / ncell)
    s = np.broadcast_to(np.eye(n), (ncell, n, n))
    j_k = kpoint_ref.coulomb(b_k, sig, w, d_k)
    k_k = kpoint_ref.exchange(b_k, sig, w, s, 0.0, d_k)
    np.testing.assert_allclose(j_k, kpoint_ref.from_supercell(jk_ref.coulomb(b_super, d_super), ncell), atol=1e-10)
    np.testing.assert_allclose(k_k, kpoint_ref.from_supercell(jk_ref.exchange(b_super, d_super), ncell), atol=1e-10)
    # Both are Hermitian per k for a Hermitian density.
    np.testing.assert_allclose(k_k, np.conj(np.swapaxes(k_k, 1, 2)), atol=1e-10)


def test_signature_subtracts_the_negative_aux_functions():
    ncell = 3
    beta, d_super = ring(4, ncell=ncell)
    naux, n = beta.shape[0], beta.shape[1]
    signs = np.array([1.0, -1.0, 1.0, -1.0])
    b_super, b_k, _ = kpoint_ref.supercell_tensor(beta)
    d_k = kpoint_ref.from_supercell(d_super, ncell)
    sig = np.broadcast_to(signs, (ncell, ncell, naux))
    w = np.full(ncell, 1.0 / ncell)
    s = np.broadcast_to(np.eye(n), (ncell, n, n))
    rows = np.arange(ncell * naux) % naux
    plus, minus = b_super[signs[rows] > 0], b_super[signs[rows] < 0]
    j_super = jk_ref.coulomb(plus, d_super) - jk_ref.coulomb(minus, d_super)
    k_super = jk_ref.exchange(plus, d_super) - jk_ref.exchange(minus, d_super)
    np.testing.assert_allclose(kpoint_ref.coulomb(b_k, sig, w, d_k), kpoint_ref.from_supercell(j_super, ncell),
                               atol=1e-10)
    np.testing.assert_allclose(kpoint_ref.exchange(b_k, sig, w, s, 0.0, d_k),
                               kpoint_ref.from_supercell(k_super, ncell), atol=1e-10)
    # The zero-transfer Coulomb builder is the same J from the diagonal blocks.
    diag = np.stack([b_k[k, k] for k in range(ncell)])
    np.testing.assert_allclose(kpoint_ref.coulomb_zero_transfer(diag, signs, w, d_k),
                               kpoint_ref.coulomb(b_k, sig, w, d_k), atol=1e-12)


def test_a_planted_mistake_fails_the_supercell_check():
    """Contracting with B^T instead of B^dagger (a dropped conjugation) breaks the equivalence."""
    ncell = 3
    beta, d_super = ring(5, ncell=ncell)
    b_super, b_k, _ = kpoint_ref.supercell_tensor(beta)
    d_k = kpoint_ref.from_supercell(d_super, ncell)
    w = np.full(ncell, 1.0 / ncell)
    half = b_k @ d_k[None, :, None, :, :]
    wrong = np.sum(np.sum(half @ np.swapaxes(b_k, -1, -2), axis=2) * w[None, :, None, None], axis=1)
    expect = kpoint_ref.from_supercell(jk_ref.exchange(b_super, d_super), ncell)
    assert np.max(np.abs(wrong - expect)) > 1e-3


def test_madelung_term_adds_the_probe_charge_correction():
    rng = np.random.default_rng(6)
    nk, naux, n = 2, 3, 4
    b = rng.standard_normal((nk, nk, naux, n, n)) + 1j * rng.standard_normal((nk, nk, naux, n, n))
    sig = np.ones((nk, nk, naux))
    w = np.full(nk, 0.5)
    a = rng.standard_normal((nk, n, n)) + 1j * rng.standard_normal((nk, n, n))
    s = a @ np.conj(np.swapaxes(a, 1, 2)) + n * np.eye(n)
    d = rng.standard_normal((nk, n, n)) + 1j * rng.standard_normal((nk, n, n))
    d = d + np.conj(np.swapaxes(d, 1, 2))
    diff = kpoint_ref.exchange(b, sig, w, s, 0.37, d) - kpoint_ref.exchange(b, sig, w, s, 0.0, d)
    for k in range(nk):
        np.testing.assert_allclose(diff[k], 0.37 * s[k] @ d[k] @ s[k], atol=1e-12)
