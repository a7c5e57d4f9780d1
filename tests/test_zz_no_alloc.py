"""The allocation-free exchange entry point against the workspace one, through the C ABI."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from hafnium_amd import synthetic
from oracle import jk_ref

pytestmark = pytest.mark.gpu


def noalloc(jk, c, signs, precision):
    import torch

    from hafnium_amd import _lib

    c = torch.as_tensor(c, dtype=torch.float64).cuda().contiguous()
    out = torch.full((jk.n_ao, jk.n_ao), float("nan"), dtype=torch.float64, device="cuda")
    s = None if signs is None else torch.as_tensor(signs, dtype=torch.float64).cuda().contiguous()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    jk._ensure_copy(_lib.precision_code(precision))
    _lib.check(jk._ctx.lib.hfjk_exchange_occ_noalloc(
        jk._ctx.handle, ctypes.c_void_p(c.data_ptr()), int(c.shape[1]),
        None if s is None else ctypes.c_void_p(s.data_ptr()), ctypes.c_void_p(out.data_ptr()),
        _lib.precision_code(precision), stream))
    return out.cpu().numpy()


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
def test_noalloc_exchange_matches_the_workspace_entry_point(precision):
    """Two calls in a row, the second smaller (fewer occupied columns on a smaller tensor), each
    against hfjk_exchange_occ; the FP64 results also against the oracle."""
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange

    cases = [(31, 90, 200, 70, None), (32, 40, 97, 9, "signed")]
    for seed, naux, nao, nocc, kind in cases:
        b, c, d = synthetic.jk_inputs(seed, naux, nao, nocc)
        jk = MI355XCoulombExchange(b)
        signs = None
        if kind == "signed":
            signs = np.where(np.arange(nocc) % 3 == 0, -1.0, 1.0)
        expect = jk.exchange_occupied(c, signs, precision=precision)
        got = noalloc(jk, c, signs, precision)
        # The same kernels over the same layout: bit for bit.
        np.testing.assert_array_equal(got, expect)
        if precision == "fp64":
            scale = np.ones(nocc) if signs is None else signs
            k_ref = jk_ref.exchange(b, (c * scale[None, :]) @ c.T)
            assert jk_ref.max_relative_error(got, k_ref) < 1e-10


def test_noalloc_rejects_problems_beyond_its_scratch():
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange
    from hafnium_amd.errors import BackendError

    jk = MI355XCoulombExchange.synthetic(3, 1500, 640)
    c = synthetic.occupied_orbitals(3, 640, 300)
    with pytest.raises(BackendError, match="bytes of scratch"):
        noalloc(jk, c, None, "fp64")


def test_noalloc_zero_occupied_columns_gives_zero():
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange

    jk = MI355XCoulombExchange.synthetic(4, 5, 33)
    got = noalloc(jk, np.zeros((33, 0)), None, "fp64")
    np.testing.assert_array_equal(got, np.zeros((33, 33)))
