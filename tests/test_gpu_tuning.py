"""Every tile schedule the autotuner can choose gives the oracle's J/K."""

from __future__ import annotations

import numpy as np
import pytest

from hafnium_amd import _lib, synthetic
from oracle import jk_ref
from tests.tolerances import BF16X3_REL_TOL, FP32_REL_TOL, FP64_REL_TOL

pytestmark = pytest.mark.gpu

TOL = {"fp64": FP64_REL_TOL, "fp32": FP32_REL_TOL, "bf16x3": BF16X3_REL_TOL}


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
def test_all_tiles_and_splits_agree_with_the_oracle(precision):
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange

    naux, nao, nocc = 13, 150, 90
    b, c, _ = synthetic.jk_inputs(77, naux, nao, nocc)
    expect = jk_ref.exchange_occupied(b, c)
    jk = MI355XCoulombExchange(b, precision=precision)
    param = _lib.TUNE_X16_BN if precision == "bf16x3" else _lib.TUNE_X_BN
    for tile in (32, 64, 128, 256):
        for splits in (8, 16, 24):
            jk.set_tuning(param, tile)
            jk.set_tuning(_lib.TUNE_SYRK_SPLITS, splits)
            k = jk.exchange_occupied(c)
            assert jk_ref.max_relative_error(k, expect) < TOL[precision], (tile, splits)
    with pytest.raises(ValueError):
        jk.set_tuning(param, 48)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("nao", [33, 150, 257])
def test_row_tiles_agree_with_the_oracle(precision, nao):
    """128-row half-transform tiles, including row blocks that stick out past the padded X rows."""
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange

    b, c, _ = synthetic.jk_inputs(91, 7, nao, min(nao, 70))
    expect = jk_ref.exchange_occupied(b, c)
    jk = MI355XCoulombExchange(b, precision=precision)
    for rows in (64, 128):
        for tile in (32, 64, 128, 256):
            jk.set_tuning(_lib.TUNE_X_BM, rows)
            jk.set_tuning(_lib.TUNE_X_BN, tile)
            k = jk.exchange_occupied(c)
            assert jk_ref.max_relative_error(k, expect) < TOL[precision], (rows, tile)
    with pytest.raises(ValueError):
        jk.set_tuning(_lib.TUNE_X_BM, 96)


def test_autotune_sets_a_valid_schedule(tmp_path):
    from hafnium_amd.autotune import autotune
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange

    b, c, _ = synthetic.jk_inputs(5, 40, 200, 60)
    jk = MI355XCoulombExchange(b, precision="bf16x3")
    cache = str(tmp_path / "tuning.json")
    result = autotune(jk, c, "bf16x3", cache_path=cache)
    assert result.x_bn in (32, 64, 128, 256) and result.syrk_splits % 8 == 0
    assert jk.get_tuning(_lib.TUNE_X16_BN) == result.x_bn
    again = autotune(MI355XCoulombExchange(b, precision="bf16x3"), c, "bf16x3", cache_path=cache)
    assert again.from_cache and (again.x_bn, again.syrk_splits) == (result.x_bn, result.syrk_splits)
    k = jk.exchange_occupied(c)
    assert jk_ref.max_relative_error(k, jk_ref.exchange_occupied(b, c)) < BF16X3_REL_TOL


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("naux,nao,nocc,splits", [
    (9, 150, 70, 8),     # last 128-row block holds 64 rows of X
    (5, 257, 33, 16),    # more splits than aux functions: empty splits
    (1, 40, 3, 8),       # one step per split (prologue repeats the last stage)
    (23, 300, 140, 8),   # several bulk iterations and a partial one
    (300, 520, 40, 8),   # several FP32 slots per split; an odd count of 128-row blocks
])
def test_syrk_kernels_agree_with_the_oracle(precision, naux, nao, nocc, splits):
    """The library's choice (BF16x3: 256 x 256 tiles), the register-staged and both direct-LDS SYRKs
    (two and one workgroups per CU) give the oracle's K."""
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange

    b, c, _ = synthetic.jk_inputs(17, naux, nao, nocc)
    expect = jk_ref.exchange_occupied(b, c)
    jk = MI355XCoulombExchange(b, precision=precision)
    jk.set_tuning(_lib.TUNE_SYRK_SPLITS, splits)
    for kind in (0, 1, 2, 3):
        jk.set_tuning(_lib.TUNE_SYRK_KIND, kind)
        k = jk.exchange_occupied(c)
        assert jk_ref.max_relative_error(k, expect) < TOL[precision], kind
    with pytest.raises(ValueError):
        jk.set_tuning(_lib.TUNE_SYRK_KIND, 4)


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
@pytest.mark.parametrize("kind", [1, 2, 3])
def test_syrk_kernels_apply_signs(kind, precision):
    """Signed columns (density form of an indefinite matrix) through every SYRK kernel."""
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange

    naux, nao, nocc = 11, 130, 45
    b, c, _ = synthetic.jk_inputs(29, naux, nao, nocc)
    signs = np.where(np.arange(nocc) % 3 == 0, -1.0, 1.0)
    expect = jk_ref.exchange(b, (c * signs) @ c.T)
    jk = MI355XCoulombExchange(b, precision=precision)
    jk.set_tuning(_lib.TUNE_SYRK_KIND, kind)
    k = jk.exchange_occupied(c, signs)
    assert jk_ref.max_relative_error(k, expect) < TOL[precision]


def test_wide_occupied_block_uses_the_register_kernel():
    """Occupied widths beyond the direct-LDS kernel's sign table still give the oracle's K."""
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange

    naux, nao, nocc = 2, 2100, 2050
    b, c, _ = synthetic.jk_inputs(3, naux, nao, nocc)
    expect = jk_ref.exchange_occupied(b, c)
    jk = MI355XCoulombExchange(b, precision="fp64")
    k = jk.exchange_occupied(c)
    assert jk_ref.max_relative_error(k, expect) < FP64_REL_TOL


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_capped_coulomb_grids_give_the_same_build(precision):
    """The grid cap of the Coulomb kernels beside the SYRK, and running them after it instead,
    change the schedule only: J and K of the combined build are bitwise equal for every choice,
    and J matches the oracle."""
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange

    naux, nao, nocc = 37, 150, 40
    b, c, d = synthetic.jk_inputs(23, naux, nao, nocc)
    jk = MI355XCoulombExchange(b, precision=precision)
    j0, k0 = jk.jk_occupied(c)
    assert jk_ref.max_relative_error(j0, jk_ref.coulomb(b, d)) < TOL[precision]
    for cap in (1, 7, 64, 1 << 20):
        jk.set_tuning(_lib.TUNE_SIDE_GROUPS, cap)
        j, k = jk.jk_occupied(c)
        np.testing.assert_array_equal(j, j0)
        np.testing.assert_array_equal(k, k0)
    jk.set_tuning(_lib.TUNE_SIDE_GROUPS, 0)
    assert jk.get_tuning(_lib.TUNE_SIDE_GROUPS) == 0
    # The Coulomb half after the SYRK on the launch stream: the same J and K.
    jk.set_tuning(_lib.TUNE_COULOMB_SERIAL, 1)
    j, k = jk.jk_occupied(c)
    np.testing.assert_array_equal(j, j0)
    np.testing.assert_array_equal(k, k0)
    jk.set_tuning(_lib.TUNE_COULOMB_SERIAL, 0)
    with pytest.raises(ValueError):
        jk.set_tuning(_lib.TUNE_COULOMB_SERIAL, 2)
