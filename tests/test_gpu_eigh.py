"""The batched lowest-eigenpair solver (hfjk_eigh_lowest_batched) against numpy.linalg.eigh.

Checked per matrix: the nvec lowest eigenvalues to 1e-12 of the spectral
norm; the nvec vectors' residuals |A x - l x| and orthonormality to 1e-10; and, where the
nvec-th and next eigenvalues are apart, the projector onto the returned
vectors equal to numpy's (the quantity the SCF density uses). Cases: random
matrices of several sizes including 1 and 2, exactly degenerate spectra
(repeated blocks), a Fock-like matrix with directions shifted by 1e3 (the
batched SCF's discarded functions), and a diagonal matrix (every reflector
empty). The sizes cover one to eight 64-row chunks per vector (n up to the
limit, 512), the register path (nvec <= 64) and the Gram-Schmidt path past
it; nvec = 0 leaves the matrices as they were.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solve(mats: np.ndarray, nvec: int, full: bool = False):
    import torch

    from hafnium_amd import _lib
    from hafnium_amd.coulomb_exchange import _Context

    ctx = _Context(0)
    dev = torch.device("cuda", 0)
    count, n, _ = mats.shape
    a = torch.from_numpy(np.ascontiguousarray(mats)).to(dev)
    w = torch.empty((count, n), dtype=torch.float64, device=dev)
    work = torch.empty(int(ctx.lib.hfjk_eigh_lowest_batched_work(n, count, nvec)), dtype=torch.float64, device=dev)
    info = torch.full((count,), 7, dtype=torch.int32, device=dev)
    p = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
    _lib.check(ctx.lib.hfjk_eigh_lowest_batched(ctx.handle, n, count, nvec, p(a), p(w), p(work), p(info),
                                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    assert (info.cpu().numpy() == 0).all()
    return w.cpu().numpy(), a.cpu().numpy() if full else a.cpu().numpy()[:, :nvec, :]


def _check(mats, nvec, w, x, projector=True):
    for i, m in enumerate(mats):
        lam, vec = np.linalg.eigh(m)
        scale = max(np.abs(lam).max(), 1e-300)
        np.testing.assert_allclose(w[i][:nvec], lam[:nvec], rtol=0, atol=1e-12 * scale)
        xi = x[i]
        resid = m @ xi.T - xi.T * w[i][:nvec]
        assert np.abs(resid).max() < 1e-10 * scale, i
        assert np.abs(xi @ xi.T - np.eye(nvec)).max() < 1e-10, i
        if projector and nvec < len(lam) and lam[nvec] - lam[nvec - 1] > 1e-6 * scale:
            want = vec[:, :nvec] @ vec[:, :nvec].T
            assert np.abs(xi.T @ xi - want).max() < 1e-10, i


@pytest.mark.parametrize("n,nvec", [(1, 1), (2, 1), (2, 2), (3, 2), (50, 12), (64, 64), (65, 64), (128, 65), (130, 40),
                                    (287, 60), (300, 300), (512, 64)])
def test_random_symmetric_matrices(n, nvec):
    rng = np.random.default_rng(n + nvec)
    g = rng.standard_normal((4, n, n))
    mats = g + g.transpose(0, 2, 1)
    w, x = _solve(mats, nvec)
    _check(mats, nvec, w, x)


def test_degenerate_and_shifted_spectra():
    rng = np.random.default_rng(5)
    n, nvec = 96, 30
    mats = []
    # Three copies of one 32 x 32 block: every eigenvalue threefold.
    b = rng.standard_normal((32, 32))
    b = b + b.T
    mats.append(np.kron(np.eye(3), b))
    # Two orthogonally mixed copies of one spectrum (degenerate pairs, dense matrix).
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.repeat(np.sort(rng.standard_normal(n // 2)), 2)
    mats.append((q * lam) @ q.T)
    # Fock-like: eigenvalues from -20 to 2, a third of the directions shifted up by 1e3.
    lam = np.concatenate([np.sort(rng.uniform(-20.0, 2.0, 64)), np.full(32, 1.0e3)])
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    mats.append((q * lam) @ q.T)
    # Diagonal, unsorted (every Householder reflector empty).
    mats.append(np.diag(rng.standard_normal(n)))
    mats = np.array(mats)
    w, x = _solve(mats, nvec)
    _check(mats, nvec, w, x)


def test_no_vectors_leaves_the_matrices():
    rng = np.random.default_rng(9)
    g = rng.standard_normal((3, 96, 96))
    mats = g + g.transpose(0, 2, 1)
    _, a = _solve(mats, 0, full=True)
    np.testing.assert_array_equal(a, mats)
