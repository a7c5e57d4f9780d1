"""k-point DF-J/K on the GPU against the k-point oracle (itself pinned by the supercell equivalence)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import jk_ref, kpoint_ref
from tests.test_kpoint_oracle import ring

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    return jk_ref.max_relative_error(np.asarray(a), np.asarray(b))


def random_hermitian(rng, nk, n, rank=None):
    a = rng.standard_normal((nk, n, n)) + 1j * rng.standard_normal((nk, n, n))
    if rank is not None:  # positive semidefinite of the given rank
        c = a[:, :, :rank]
        return c @ np.conj(np.swapaxes(c, 1, 2))
    return 0.5 * (a + np.conj(np.swapaxes(a, 1, 2)))


@pytest.mark.parametrize("ncell", [1, 3, 4])
def test_supercell_ring_matches_the_oracle(ncell):
    from hafnium_amd.kpoint import KPointMI355XCoulombExchange

    beta, d_super = ring(10 + ncell, naux=6, nao=5, ncell=ncell)
    _, b_k, _ = kpoint_ref.supercell_tensor(beta)
    d_k = kpoint_ref.from_supercell(d_super, ncell)
    sig = np.ones(b_k.shape[:3])
    sig[:, :, 1] = -1.0
    w = np.full(ncell, 1.0 / ncell)
    s = np.broadcast_to(np.eye(5), (ncell, 5, 5)).astype(np.complex128)
    jk = KPointMI355XCoulombExchange(b_k, sig, w, s, 0.0)
    assert rel(jk.coulomb(d_k), kpoint_ref.coulomb(b_k, sig, w, d_k)) < TOL
    assert rel(jk.exchange(d_k), kpoint_ref.exchange(b_k, sig, w, s, 0.0, d_k)) < TOL


@pytest.mark.parametrize("nk,naux,nao,rank", [(2, 40, 20, None), (3, 25, 33, 4), (1, 70, 64, None), (2, 9, 70, 12)])
def test_general_blocks_match_the_oracle(nk, naux, nao, rank):
    """Random complex transfer blocks, mixed signatures, uneven weights, the Madelung term, and
    indefinite or low-rank Hermitian densities."""
    from hafnium_amd.kpoint import KPointMI355XCoulombExchange

    rng = np.random.default_rng(100 * nk + nao)
    b = rng.standard_normal((nk, nk, naux, nao, nao)) + 1j * rng.standard_normal((nk, nk, naux, nao, nao))
    b *= np.exp(-np.abs(np.subtract.outer(np.arange(nao), np.arange(nao))) / 10.0)
    sig = np.where(rng.random((nk, nk, naux)) < 0.2, -1.0, 1.0)
    w = rng.random(nk) + 0.1
    w /= w.sum()
    s = random_hermitian(rng, nk, nao, rank=nao) + nao * np.eye(nao)
    d = random_hermitian(rng, nk, nao, rank=rank)
    jk = KPointMI355XCoulombExchange(b, sig, w, s, 0.29)
    assert rel(jk.coulomb(d), kpoint_ref.coulomb(b, sig, w, d)) < TOL
    assert rel(jk.exchange(d), kpoint_ref.exchange(b, sig, w, s, 0.29, d)) < TOL


def test_zero_transfer_builder_and_device_tensors():
    import torch

    from hafnium_amd.errors import InputError
    from hafnium_amd.kpoint import KPointMI355XCoulomb

    rng = np.random.default_rng(7)
    nk, naux, nao = 3, 30, 17
    b = rng.standard_normal((nk, naux, nao, nao)) + 1j * rng.standard_normal((nk, naux, nao, nao))
    sig = np.where(rng.random(naux) < 0.3, -1.0, 1.0)
    w = np.full(nk, 1.0 / nk)
    d = random_hermitian(rng, nk, nao)
    jk = KPointMI355XCoulomb(b, sig, w)
    got = jk.coulomb(torch.from_numpy(d).cuda())
    assert isinstance(got, torch.Tensor)
    assert rel(got.cpu().numpy(), kpoint_ref.coulomb_zero_transfer(b, sig, w, d)) < TOL
    with pytest.raises(InputError, match="Coulomb term alone"):
        jk.exchange(d)


def test_malformed_inputs_are_rejected():
    from hafnium_amd.errors import InputError
    from hafnium_amd.kpoint import KPointMI355XCoulombExchange

    b = np.zeros((2, 2, 3, 4, 4), dtype=complex)
    with pytest.raises(InputError, match="kpt, kpt, aux, ao, ao"):
        KPointMI355XCoulombExchange(b[0], np.ones((2, 3)), np.ones(2), np.zeros((2, 4, 4)), 0.0)
    with pytest.raises(InputError, match="signature"):
        KPointMI355XCoulombExchange(b, np.ones((2, 3)), np.ones(2), np.zeros((2, 4, 4)), 0.0)
    with pytest.raises(InputError, match="weights"):
        KPointMI355XCoulombExchange(b, np.ones((2, 2, 3)), np.ones(3), np.zeros((2, 4, 4)), 0.0)
    jk = KPointMI355XCoulombExchange(b, np.ones((2, 2, 3)), np.full(2, 0.5), np.zeros((2, 4, 4)), 0.0)
    with pytest.raises(InputError, match="not Hermitian"):
        jk.exchange(np.ones((2, 4, 4)) + 1j * np.triu(np.ones((2, 4, 4))))
    with pytest.raises(InputError, match="shape"):
        jk.coulomb(np.zeros((2, 3, 3)))
