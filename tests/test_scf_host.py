"""Host logic of the GPU SCF driver: DIIS solve, degenerate-group filling, options."""

from __future__ import annotations

import numpy as np
import pytest

from hafnium_amd import scf
from hafnium_amd.errors import InputError
from oracle import scf_ref
from tests.tolerances import STRUCTURE_TOL


def test_diis_solve_matches_the_oracle_and_the_linear_model():
    # Errors e_i = x_i - 2 on a one-dimensional linear model: c = (1/2, 1/2).
    ov = 2.0 * np.array([[1.0, -1.0], [-1.0, 1.0]])
    np.testing.assert_allclose(scf.solve_diis(ov), [0.5, 0.5], atol=STRUCTURE_TOL)
    rng = np.random.default_rng(0)
    e = rng.standard_normal((5, 30))
    ov = e @ e.T
    np.testing.assert_allclose(scf.solve_diis(ov), scf_ref.solve_diis(ov), atol=1e-12)
    # Singular (repeated error) and all-zero systems stay finite.
    dup = np.array([[1.0, 1.0, 1.0], [1.0, 1.0, 1.0], [1.0, 1.0, 1.0]])
    assert np.all(np.isfinite(scf.solve_diis(dup)))
    np.testing.assert_allclose(scf.solve_diis(np.zeros((2, 2))), [0.5, 0.5], atol=STRUCTURE_TOL)


def test_degenerate_groups_share_electrons():
    occ = scf.fill_degenerate_groups(np.array([-2.0, -1.0, -1.0, 0.5]), 2.0)
    np.testing.assert_allclose(occ, [1.0, 0.5, 0.5, 0.0])
    np.testing.assert_allclose(occ, scf_ref.fill_degenerate_groups(np.array([-2.0, -1.0, -1.0, 0.5]), 2.0))
    with pytest.raises(InputError):
        scf.fill_degenerate_groups(np.array([0.0, 1.0]), 3.0)


def test_options_are_validated():
    # The words the reference's own option tests match on (tests/unit/test_scf.py:275-286).
    for kwargs, word in (({"max_iterations": 0}, "iteration"), ({"energy_tolerance": 0.0}, "positive"),
                         ({"diis_size": 0}, "DIIS"), ({"level_shift": -0.1}, "level shift"),
                         ({"damping": 1.0}, "damping"), ({"overlap_threshold": 0.0}, "overlap threshold")):
        with pytest.raises(InputError, match=word):
            scf.ScfOptions(**kwargs)
    with pytest.raises(InputError):
        scf.ScfOptions(max_iterations=0)
    with pytest.raises(InputError):
        scf.ScfOptions(precision="fp16")
    with pytest.raises(InputError):
        scf.ScfOptions(damping=1.0)
    terms = scf.EnergyTerms(9.0, -120.0, 45.0, -9.0)
    assert terms.total == pytest.approx(9.0 - 120.0 + 45.0 - 9.0)


def test_mixed_phase_stall_detection():
    # Geometric convergence never stalls.
    assert not scf.mixed_phase_stalled([1e-1, 3e-2, 1e-2, 3e-3, 1e-3], 2, 0.7)
    # A floor near 1e-4 (the reduced-precision J/K noise) does.
    assert scf.mixed_phase_stalled([1e-1, 1e-2, 1e-3, 1.2e-4, 1.1e-4, 1.3e-4], 2, 0.7)
    # Too few iterations to tell.
    assert not scf.mixed_phase_stalled([1e-1, 2e-1, 3e-1], 2, 0.7)
    # One bad iteration followed by progress is not a stall.
    assert not scf.mixed_phase_stalled([1e-1, 1e-2, 2e-2, 5e-3], 2, 0.7)
    with pytest.raises(InputError):
        scf.ScfOptions(stall_ratio=0.0)


def test_balanced_groups_deal_by_cost():
    from hafnium_amd.batch_scf import balanced_groups

    groups = balanced_groups([48, 288, 96, 240, 144, 192], 2)
    assert sorted(i for g in groups for i in g) == list(range(6))
    cost = [sum([48, 288, 96, 240, 144, 192][i] ** 4 for i in g) for g in groups]
    assert max(cost) / min(cost) < 1.6
    assert balanced_groups([10], 3) == [[0]]


def test_switch_point_is_the_threshold_raised_to_the_precision_floor():
    from hafnium_amd.errors import InputError
    from hafnium_amd.scf import ScfOptions

    opts = ScfOptions()
    assert opts.switch_at("fp64") == opts.switch_threshold == 1e-5
    assert opts.switch_at("fp32") == 1e-4 and opts.switch_at("bf16x3") == 1e-3
    low = ScfOptions(switch_floor_fp32=0.0, switch_floor_bf16x3=1e-6)
    assert low.switch_at("fp32") == 1e-5 and low.switch_at("bf16x3") == 1e-5
    assert ScfOptions(switch_threshold=2e-3).switch_at("bf16x3") == 2e-3
    with pytest.raises(InputError):
        ScfOptions(switch_floor_fp32=-1.0)
