"""Parity of the HIP DF-J/K path with the oracle, through the C ABI, on an MI355X."""

from __future__ import annotations

import numpy as np
import pytest

from hafnium_amd import synthetic
from hafnium_amd.errors import BackendError, InputError
from oracle import jk_ref, layout_ref, scf_ref
from tests.tolerances import BF16X3_REL_TOL, ENERGY_TOL, FP32_REL_TOL, FP64_REL_TOL

pytestmark = pytest.mark.gpu

TOL = {"fp64": FP64_REL_TOL, "fp32": FP32_REL_TOL, "bf16x3": BF16X3_REL_TOL}


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available(), "the gpu tests need an MI355X"
    return torch


def builder(b, **kw):
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange

    return MI355XCoulombExchange(b, **kw)


def rel(a, b):
    return jk_ref.max_relative_error(np.asarray(a), np.asarray(b))


def test_generator_matches_numpy_bit_for_bit(torch):
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange

    for seed, naux, nao, q0 in ((0, 3, 70, 0), (9, 2, 33, 5)):
        jk = MI355XCoulombExchange.synthetic(seed, naux, nao, q0=q0)
        packed = jk.packed_tensor("fp64").view(naux, -1).cpu().numpy()
        expect = layout_ref.pack(synthetic.tensor_slice_full(seed, nao, q0, q0 + naux))
        np.testing.assert_array_equal(packed, expect)


def test_normal_matrix_matches_numpy(torch):
    import ctypes

    from hafnium_amd import _lib
    from hafnium_amd.coulomb_exchange import _Context

    ctx = _Context(0)
    out = torch.empty((37, 5), dtype=torch.float64, device="cuda")
    _lib.check(ctx.lib.hfjk_normal_matrix(ctx.handle, 3, 37, 5, synthetic.TAG_ORBITALS,
                                          ctypes.c_void_p(out.data_ptr()), None))
    k, i = np.meshgrid(np.arange(37), np.arange(5), indexing="ij")
    expect = synthetic.normal_from_counters(k, i, 0xFFFFFFFF, synthetic.TAG_ORBITALS, 3)
    np.testing.assert_array_equal(out.cpu().numpy(), expect)


def test_pack_matches_the_layout(torch):
    b = synthetic.tensor_full(2, 4, 45)
    jk = builder(b)
    np.testing.assert_array_equal(jk.packed_tensor().view(4, -1).cpu().numpy(), layout_ref.pack(b))


def test_fp64_matches_golden_c0(torch, synthetic_c0):
    g = synthetic_c0
    jk = builder(synthetic.tensor_full(int(g["seed"]), int(g["naux"]), int(g["nao"])))
    d = g["density"]
    j = jk.coulomb(d)
    k_dens = jk.exchange(d)
    k_occ = jk.exchange_occupied(g["orbitals"])
    assert rel(j, g["coulomb"]) < FP64_REL_TOL
    assert rel(k_dens, g["exchange"]) < FP64_REL_TOL
    assert rel(k_occ, g["exchange"]) < FP64_REL_TOL


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("naux,nao,nocc", [(1, 1, 1), (7, 33, 1), (5, 64, 33), (11, 65, 40), (3, 130, 130),
                                           (17, 97, 70), (4, 200, 100), (3, 300, 300), (9, 129, 17)])
def test_ragged_shapes_match_the_oracle(torch, precision, naux, nao, nocc):
    # The occupied widths cover every column-tile shape of the half-transform
    # (32, 64, 128, 256 and two 256-wide tiles) and partial 128-row SYRK tiles.
    seed = 100 + nao
    b, c, d = synthetic.jk_inputs(seed, naux, nao, nocc)
    jk = builder(b, precision=precision)
    tol = TOL[precision]
    assert rel(jk.coulomb(d), jk_ref.coulomb(b, d)) < tol
    k_ref = jk_ref.exchange(b, d)
    assert rel(jk.exchange_occupied(c), k_ref) < tol
    assert rel(jk.exchange(d), k_ref) < tol


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("naux,nao,nocc", [(1, 1, 1), (2, 33, 3), (11, 65, 40), (3, 130, 130), (17, 97, 70),
                                           (3, 300, 300), (9, 129, 17)])
def test_combined_build_matches_the_oracle(torch, precision, naux, nao, nocc):
    """J and K of D = C C^T in one call (charges formed from the half-transformed X)."""
    b, c, d = synthetic.jk_inputs(200 + nao, naux, nao, nocc)
    jk = builder(b, precision=precision)
    j, k = jk.jk_occupied(c)
    assert rel(j, jk_ref.coulomb(b, d)) < TOL[precision]
    assert rel(k, jk_ref.exchange(b, d)) < TOL[precision]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_combined_build_over_every_tile_choice(torch, precision):
    """The charges from X give the same J for every half-transform tiling."""
    from hafnium_amd import _lib

    b, c, d = synthetic.jk_inputs(41, 6, 150, 90)
    jk = builder(b, precision=precision)
    j_ref = jk_ref.coulomb(b, d)
    for rows in (64, 128):
        for tile in (32, 64, 128, 256):
            jk.set_tuning(_lib.TUNE_X_BM, rows)
            jk.set_tuning(_lib.TUNE_X_BN, tile)
            j, _ = jk.jk_occupied(c)
            assert rel(j, j_ref) < TOL[precision], (rows, tile)


def test_combined_build_is_deterministic_and_equals_the_separate_calls(torch):
    b, c, d = synthetic.jk_inputs(13, 40, 90, 20)
    jk = builder(b)
    j1, k1 = jk.jk_occupied(c)
    j2, k2 = jk.jk_occupied(c)
    np.testing.assert_array_equal(j1, j2)
    np.testing.assert_array_equal(k1, k2)
    np.testing.assert_array_equal(k1, jk.exchange_occupied(c))
    assert rel(j1, jk.coulomb(d)) < 1e-13


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_indefinite_and_nonsymmetric_densities(torch, precision):
    rng = np.random.default_rng(5)
    b = synthetic.tensor_full(6, 20, 50)
    a = rng.standard_normal((50, 50))
    sym = a + a.T  # indefinite, full rank
    jk = builder(b, precision=precision)
    assert rel(jk.exchange(sym), jk_ref.exchange(b, sym)) < TOL[precision]
    # J depends on the symmetric part only, as the oracle's full contraction.
    assert rel(jk.coulomb(a), jk_ref.coulomb(b, a)) < TOL[precision]
    # A non-symmetric density: the Hermitian factorisation of S + i A and two SYRKs, in every precision.
    assert rel(jk.exchange(a), jk_ref.exchange(b, a)) < TOL[precision]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_c_abi_density_exchange_matches_the_two_step_path(torch, precision):
    """hfjk_exchange_density (one call, full-rank workspace) against the factor + occupied-form
    path of MI355XCoulombExchange.exchange (rank-sized workspace) and the oracle."""
    import ctypes

    from hafnium_amd import _lib

    rng = np.random.default_rng(9)
    b = synthetic.tensor_full(12, 14, 70)
    jk = builder(b, precision=precision)
    a = rng.standard_normal((70, 70))
    c = rng.standard_normal((70, 9))
    prec = _lib.precision_code(precision)
    jk._ensure_copy(prec)  # the FP32 copy the direct call needs
    w = np.r_[np.ones(4), np.full(3, 2 / 3)]  # fractional (averaged) occupations
    q, _ = np.linalg.qr(rng.standard_normal((70, 7)))
    # Indefinite (dsyevd), low-rank and full-rank positive semidefinite and fractionally occupied
    # (pivoted Cholesky), zero.
    for d in (a + a.T, c @ c.T, a @ a.T + np.eye(70), (q * w) @ q.T, np.zeros((70, 70))):
        dt = torch.from_numpy(d).cuda()
        out = torch.empty((70, 70), dtype=torch.float64, device="cuda")
        ws, nbytes = jk._workspace_for(70, prec)
        _lib.check(jk._ctx.lib.hfjk_exchange_density(jk._ctx.handle, ctypes.c_void_p(dt.data_ptr()),
                                                     ctypes.c_void_p(out.data_ptr()), prec, ws, nbytes,
                                                     jk._stream()))
        two_step = jk.exchange(d)
        ref = jk_ref.exchange(b, d)
        if np.any(d):
            assert rel(out.cpu().numpy(), ref) < TOL[precision]
            assert rel(two_step, ref) < TOL[precision]
        else:
            assert not np.any(out.cpu().numpy()) and not np.any(two_step)
    dt = torch.from_numpy(a).cuda()
    out = torch.empty((70, 70), dtype=torch.float64, device="cuda")
    ws, nbytes = jk._workspace_for(70, prec)
    _lib.check(jk._ctx.lib.hfjk_exchange_density(jk._ctx.handle, ctypes.c_void_p(dt.data_ptr()),
                                                 ctypes.c_void_p(out.data_ptr()), prec, ws, nbytes,
                                                 jk._stream()))
    assert rel(out.cpu().numpy(), jk_ref.exchange(b, a)) < TOL[precision]


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("naux,nao,chunk,full", [(5, 1, 0, False), (7, 33, 3, False), (9, 70, 0, True),
                                                 (6, 70, 4, True), (4, 130, 1, False)])
def test_nonsymmetric_and_complex_densities_take_the_density_form(torch, naux, nao, chunk, full, precision):
    """Densities of the reference's tests (tests/unit/test_coulomb_exchange.py:45-105): random
    Hermitian, real non-symmetric, and the off-diagonal spin block of a Hermitian spinor density,
    over one or several aux chunks, with and without the full copy of B, in every precision."""
    from hafnium_amd import _lib

    rng = np.random.default_rng(nao + naux)
    b = synthetic.tensor_full(11 + nao, naux, nao)
    jk = builder(b, precision=precision)
    if full:
        jk.expand_full(precision)
    jk.set_tuning(_lib.TUNE_GENERAL_CHUNK, chunk)
    tol = TOL[precision]
    nonsym = rng.standard_normal((nao, nao))
    assert rel(jk.exchange(nonsym), jk_ref.exchange(b, nonsym)) < tol
    z = rng.standard_normal((2 * nao, 2 * nao)) + 1j * rng.standard_normal((2 * nao, 2 * nao))
    spinor = z + z.conj().T
    herm, off = spinor[:nao, :nao], spinor[:nao, nao:]
    k = jk.exchange(herm)
    j = jk.coulomb(herm)
    assert k.dtype == np.complex128 and j.dtype == np.complex128
    assert rel(k, jk_ref.exchange(b, herm)) < tol
    assert rel(j, jk_ref.coulomb(b, herm)) < tol
    # Both matrices are Hermitian for a Hermitian density; J is real.
    assert rel(k, k.conj().T) < (1e-14 if precision == "fp64" else 10 * tol)
    assert np.max(np.abs(j.imag)) <= (1e-13 if precision == "fp64" else 10 * tol) * np.max(np.abs(j.real))
    assert rel(jk.exchange(off), jk_ref.exchange(b, off)) < tol
    if precision != "fp64":
        return
    # Torch in, torch out, on the device.
    kt = jk.exchange(torch.from_numpy(off).cuda())
    assert kt.is_cuda and kt.dtype == torch.complex128
    np.testing.assert_array_equal(kt.cpu().numpy(), jk.exchange(off))
    # A complex density with a zero imaginary part gives the real result.
    sym = nonsym + nonsym.T
    np.testing.assert_array_equal(jk.exchange(sym.astype(np.complex128)).real, jk.exchange(sym))


def test_signed_orbitals_give_the_difference_of_exchanges(torch):
    b = synthetic.tensor_full(8, 9, 40)
    c = synthetic.occupied_orbitals(8, 40, 10)
    signs = np.array([1.0] * 6 + [-1.0] * 4)
    jk = builder(b)
    expect = jk_ref.exchange_occupied(b, c[:, :6]) - jk_ref.exchange_occupied(b, c[:, 6:])
    assert rel(jk.exchange_occupied(c, signs), expect) < FP64_REL_TOL


def test_fp64_is_deterministic(torch):
    b, c, d = synthetic.jk_inputs(12, 40, 90, 20)
    jk = builder(b)
    j1, k1 = jk.coulomb(d), jk.exchange_occupied(c)
    j2, k2 = jk.coulomb(d), jk.exchange_occupied(c)
    np.testing.assert_array_equal(j1, j2)
    np.testing.assert_array_equal(k1, k2)


def test_full_width_slice_of_the_large_configuration(torch):
    """nao=1500, nocc=250 (the large configuration's shapes) on a slice of 3 aux functions."""
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange

    nao, nocc, q0, naux = 1500, 250, 4321, 3
    b = synthetic.tensor_slice_full(1, nao, q0, q0 + naux)
    c = synthetic.occupied_orbitals(1, nao, nocc)
    d = c @ c.T
    for precision in ("fp64", "fp32", "bf16x3"):
        jk = MI355XCoulombExchange.synthetic(1, naux, nao, q0=q0, precision=precision)
        assert rel(jk.coulomb(d), jk_ref.coulomb(b, d)) < TOL[precision]
        assert rel(jk.exchange_occupied(c), jk_ref.exchange_occupied(b, c)) < TOL[precision]


def test_water_df_rhf_energy_through_the_gpu_path(torch, water):
    overlap = water["overlap"]
    core = water["kinetic"] + water["nuclear_attraction"]
    e_nuc = scf_ref.nuclear_repulsion(water["charges"], water["positions_bohr"])
    b = scf_ref.df_tensor(water["metric"], water["three_center"], "cholesky")
    jk = builder(b)
    result = scf_ref.rhf(overlap, core, jk.coulomb, jk.exchange, 5, e_nuc)
    assert result.energy == pytest.approx(float(water["df_rhf_energy"]), abs=ENERGY_TOL)


def test_malformed_inputs_are_rejected(torch):
    with pytest.raises(InputError, match="aux, ao, ao"):
        builder(np.zeros((3, 2, 4)))
    with pytest.raises(BackendError, match="complex"):
        builder(np.zeros((2, 3, 3), dtype=complex))
    asym = np.random.default_rng(0).standard_normal((2, 5, 5))
    with pytest.raises(InputError, match="not symmetric"):
        builder(asym)
    jk = builder(synthetic.tensor_full(1, 2, 6))
    with pytest.raises(InputError, match="atomic orbitals"):
        jk.coulomb(np.eye(7))
    with pytest.raises(BackendError, match="complex"):
        jk.exchange_occupied(np.eye(6, dtype=complex))
    with pytest.raises(InputError, match="atomic orbitals"):
        jk.exchange(np.eye(5, dtype=complex))
    with pytest.raises(InputError):
        jk.exchange_occupied(np.ones((5, 2)))


def test_mixed_copies_of_a_tensor_beyond_two_to_the_32_elements(torch):
    """The FP32 and BF16 copies cover every element when naux * packed > 2**32."""
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange

    nao, naux = 1500, 3800  # 3800 * 1155072 = 4.39e9 packed elements
    jk = MI355XCoulombExchange.synthetic(2, naux, nao)
    b64 = jk.packed_tensor("fp64")
    assert b64.numel() > 2**32
    idx = torch.tensor([0, 2**32 - 1, 2**32, 2**32 + 12345, b64.numel() - 1], device=b64.device)
    b32 = jk.packed_tensor("fp32")
    np.testing.assert_array_equal(b32[idx].cpu().numpy(), b64[idx].float().cpu().numpy())
    del b32
    jk._b32 = None
    b16 = jk.packed_tensor("bf16x3")
    count = b64.numel()
    hi = b16[idx].cpu().numpy().astype(np.uint16).astype(np.uint32) << 16
    lo = b16[idx + count].cpu().numpy().astype(np.uint16).astype(np.uint32) << 16
    value = hi.view(np.float32).astype(np.float64) + lo.view(np.float32).astype(np.float64)
    x = b64[idx].cpu().numpy()
    assert np.all(np.abs(value - x) <= 2.0**-15 * np.abs(x) + 1e-300)


def bf16_planes(x):
    """The hi/lo bf16 split of float64 values (hi = bf16(fp32(x)), lo = bf16(fp32(x) - hi)), as uint16."""
    import torch as _t

    f = _t.from_numpy(np.ascontiguousarray(x, dtype=np.float32))
    hi = f.to(_t.bfloat16)
    lo = (f - hi.to(_t.float32)).to(_t.bfloat16)
    return hi.view(_t.int16).numpy().view(np.uint16), lo.view(_t.int16).numpy().view(np.uint16)


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("naux,nao,nocc", [(1, 1, 1), (3, 33, 5), (7, 97, 70), (5, 150, 140), (4, 257, 33),
                                           (2, 300, 300), (3, 600, 200)])
def test_full_copy_half_transform_matches_the_oracle(torch, precision, naux, nao, nocc):
    """Half-transform on the direct-LDS ring from the full copy of B: X rows past npad,
    one to three column blocks, partial row blocks (BF16x3 past 128 columns: 256 x 256 tiles,
    with partial row and column tiles at nao 300 and 600)."""
    from hafnium_amd import _lib

    b, c, d = synthetic.jk_inputs(300 + nao, naux, nao, nocc)
    jk = builder(b, precision=precision)
    k_packed = jk.exchange_occupied(c)
    jk.expand_full(precision)
    npad = -(-nao // 32) * 32
    expect_full = np.zeros((naux, npad, npad))
    expect_full[:, :nao, :nao] = b
    if precision == "bf16x3":
        # Each 32-column block of a row: the 32 hi parts, then the same 32 lo parts.
        full = jk._full[_lib.BF16X3].view(naux, npad, npad // 32, 2, 32).cpu().numpy().view(np.uint16)
        hi, lo = bf16_planes(expect_full)
        np.testing.assert_array_equal(full[:, :, :, 0, :].reshape(naux, npad, npad), hi)
        np.testing.assert_array_equal(full[:, :, :, 1, :].reshape(naux, npad, npad), lo)
    else:
        full = jk._full[_lib.precision_code(precision)].view(naux, npad, npad).cpu().numpy()
        np.testing.assert_array_equal(full, expect_full.astype(full.dtype))
    k = jk.exchange_occupied(c)
    j, k2 = jk.jk_occupied(c)
    tol = TOL[precision]
    assert rel(k, jk_ref.exchange(b, d)) < tol
    assert rel(k2, jk_ref.exchange(b, d)) < tol
    assert rel(j, jk_ref.coulomb(b, d)) < tol
    # Signed columns from the same copy (BF16x3: X in hi/lo planes for the signed SYRK, where the
    # unsigned build above interleaves them past 128 columns).
    signs = np.where(np.arange(nocc) % 3 == 1, -1.0, 1.0)
    assert rel(jk.exchange_occupied(c, signs), jk_ref.exchange(b, (c * signs) @ c.T)) < tol
    assert rel(k, k_packed) < {"fp64": 1e-13, "fp32": 1e-5, "bf16x3": 1e-4}[precision]
    if precision == "bf16x3":
        # The 128-row ring from the same copy (a column tile below 256) gives the same K.
        jk.set_tuning(_lib.TUNE_X16_BN, 128)
        assert rel(jk.exchange_occupied(c), k) < 1e-6
        jk.set_tuning(_lib.TUNE_X16_BN, 0)
    jk.set_tuning(_lib.TUNE_X_FULL, 1)
    np.testing.assert_array_equal(jk.exchange_occupied(c), k_packed)
    jk.set_tuning(_lib.TUNE_X_FULL, 0)
    with pytest.raises(ValueError):
        jk.set_tuning(_lib.TUNE_X_FULL, 2)


def test_full_copy_survives_rebinding_and_can_be_dropped(torch):
    b, c, d = synthetic.jk_inputs(17, 6, 90, 40)
    jk = builder(b)
    jk.expand_full()
    k_full = jk.exchange_occupied(c)
    jk.exchange_occupied(c, precision="fp32")  # the FP32 copy rebinds the tensors
    np.testing.assert_array_equal(jk.exchange_occupied(c), k_full)
    with pytest.raises(InputError):
        jk.expand_full("fp16")
    jk.drop_full()
    assert rel(jk.exchange_occupied(c), k_full) < 1e-13
    assert rel(k_full, jk_ref.exchange(b, d)) < FP64_REL_TOL
