"""Batched DIIS coefficients on the GPU against the host solve (the reference's solve_diis)."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _host_coefficients(ov, valid, m):
    from hafnium_amd.scf import solve_diis

    out = np.zeros(m)
    if len(valid) == 1:
        out[valid[0]] = 1.0
    else:
        out[valid] = solve_diis(ov[np.ix_(valid, valid)])
    return out


@pytest.mark.parametrize("valid", [[3], [0, 1], [5, 6, 7, 0, 1, 2], [2, 3, 4, 5, 6, 7, 0, 1]])
def test_device_diis_matches_the_host_solve(valid):
    import torch

    from hafnium_amd import _lib
    from hafnium_amd.coulomb_exchange import _Context

    ctx = _Context(0)
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(len(valid))
    count, m, length, newest = 7, 8, 40, valid[-1]
    errors = rng.standard_normal((count, m, length)) * np.logspace(0, -2, m)[None, :, None]
    # Instance 3: two equal errors (a singular overlap matrix, its null space dropped);
    # instance 4: all errors zero (nothing to scale).
    errors[3, valid[0]] = errors[3, valid[-1]]
    errors[4] = 0.0
    ov_full = 2.0 * np.einsum("iak,ibk->iab", errors, errors)
    # The device keeps ov across calls; here it starts with every row but the newest.
    ov = ov_full.copy()
    ov[:, newest, :] = 7.0
    ov[:, :, newest] = 7.0
    dots = np.einsum("ik,iak->ia", errors[:, newest], errors)
    ov_dev = torch.from_numpy(ov).to(dev)
    dots_dev = torch.from_numpy(dots).to(dev)
    coef_dev = torch.full((count, m), 5.0, dtype=torch.float64, device=dev)
    slots = (ctypes.c_int * len(valid))(*valid)
    _lib.check(ctx.lib.hfjk_diis_batched(ctx.handle, count, m, ctypes.c_void_p(ov_dev.data_ptr()),
                                         ctypes.c_void_p(dots_dev.data_ptr()), newest, slots, len(valid),
                                         ctypes.c_void_p(coef_dev.data_ptr()),
                                         ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    np.testing.assert_allclose(ov_dev.cpu().numpy(), ov_full, rtol=1e-13, atol=1e-300)
    got = coef_dev.cpu().numpy()
    for i in range(count):
        want = _host_coefficients(ov_full[i], valid, m)
        np.testing.assert_allclose(got[i], want, rtol=1e-8, atol=1e-10 * max(1.0, np.abs(want).max()))
        assert got[i][valid].sum() == pytest.approx(1.0, abs=1e-9) or i == 4
        assert np.all(got[i][[a for a in range(m) if a not in valid]] == 0.0)


def test_device_diis_rejects_bad_slot_lists():
    import torch

    from hafnium_amd import _lib
    from hafnium_amd.coulomb_exchange import _Context

    ctx = _Context(0)
    buf = torch.zeros(64, dtype=torch.float64, device="cuda")
    p = ctypes.c_void_p(buf.data_ptr())
    for valid, nslots in (([1, 1], 4), ([4], 4), ([0], 17)):
        slots = (ctypes.c_int * len(valid))(*valid)
        assert ctx.lib.hfjk_diis_batched(ctx.handle, 1, nslots, p, p, 0, slots, len(valid), p, None) == _lib.ERR_INPUT


def test_device_extrapolation_combines_the_stored_matrices():
    import torch

    from hafnium_amd import _lib
    from hafnium_amd.coulomb_exchange import _Context

    ctx = _Context(0)
    rng = np.random.default_rng(3)
    nslots, count, length = 5, 3, 1000
    hist = rng.standard_normal((nslots, count, length))
    coef = rng.standard_normal((count, nslots))
    coef[1, 2] = 0.0
    hist[2, 1] = np.nan  # a slot with a zero coefficient is not read
    dev = torch.device("cuda", 0)
    h, c = torch.from_numpy(hist).to(dev), torch.from_numpy(coef).to(dev)
    out = torch.empty((count, length), dtype=torch.float64, device=dev)
    _lib.check(ctx.lib.hfjk_diis_combine_batched(ctx.handle, count, length, nslots, ctypes.c_void_p(h.data_ptr()),
                                                 ctypes.c_void_p(c.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    want = np.einsum("ia,aie->ie", coef, np.nan_to_num(hist))
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=1e-13, atol=1e-13)
