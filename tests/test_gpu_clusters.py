"""DF-RHF of water clusters on the GPU: FP64 against the oracle, mixed against FP64, the atomic guess."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import scf_ref
from tests.tolerances import ENERGY_TOL

pytestmark = pytest.mark.gpu

MIXED_ENERGY_TOL = 1e-8  # the mixed-precision target: within 1e-8 Eh of plain FP64


@pytest.fixture(scope="module")
def dimer():
    from hafnium_amd.clusters import build_system
    from hafnium_amd.molecule import water_cluster

    return build_system(water_cluster(2, seed=11))


def test_fp64_cluster_energy_matches_the_oracle(dimer):
    from hafnium_amd import integrals
    from hafnium_amd.clusters import run_cluster
    from hafnium_amd.molecule import BasisOnAtoms

    result, _ = run_cluster(2, 11, system=dimer, energy_tolerance=1e-10)
    mol = dimer.molecule
    orb = BasisOnAtoms("cc-pvdz", mol)
    aux = BasisOnAtoms("def2-universal-jkfit", mol)
    s, t, v = integrals.one_electron(orb, mol)
    b = scf_ref.df_tensor(integrals.metric(aux), integrals.three_center(aux, orb), "cholesky")
    j_of, k_of = scf_ref.fitted_operators(b)
    ref = scf_ref.rhf(s, t + v, j_of, k_of, mol.n_electrons // 2, mol.nuclear_repulsion())
    assert ref.converged
    assert result.energy == pytest.approx(ref.energy, abs=ENERGY_TOL * 10)


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
def test_mixed_precision_reaches_the_fp64_energy(dimer, precision):
    from hafnium_amd.clusters import run_cluster

    plain, _ = run_cluster(2, 11, system=dimer, precision="fp64")
    mixed, _ = run_cluster(2, 11, system=dimer, precision=precision)
    assert mixed.converged
    assert abs(mixed.energy - plain.energy) < MIXED_ENERGY_TOL
    assert any(row.precision == precision for row in mixed.history)
    assert mixed.history[-1].precision == "fp64"


def test_atomic_density_guess_is_a_valid_start(dimer):
    from hafnium_amd.clusters import atomic_density_guess, run_cluster

    guess = atomic_density_guess(dimer.molecule)
    s = dimer.hamiltonian.overlap.cpu().numpy()
    # Electron count of the superposition: tr(D S) = n_alpha (each atom's share).
    assert float(np.trace(guess @ s)) == pytest.approx(dimer.molecule.n_electrons / 2, abs=0.05)
    core, _ = run_cluster(2, 11, system=dimer)
    sad, _ = run_cluster(2, 11, system=dimer, guess=guess)
    assert sad.energy == pytest.approx(core.energy, abs=MIXED_ENERGY_TOL)
    assert sad.n_iterations <= core.n_iterations
