"""Exact four-centre contractions on the GPU and the spherical-atom density guess built on them."""

from __future__ import annotations

import numpy as np
import pytest

from hafnium_amd.errors import InputError
from oracle import scf_ref
from tests.tolerances import ENERGY_TOL, MO_ENERGY_TOL, STRUCTURE_TOL

pytestmark = pytest.mark.gpu


def test_exact_builder_matches_the_looped_contractions(water):
    """J and K of the factorised stored integrals against the oracle's einsum contractions (the
    reference's ExactCoulombExchange, coulomb_exchange.py:113-121) for the densities of the
    reference's tests: symmetric, non-symmetric and complex Hermitian."""
    from hafnium_amd.density_fitting import exact_coulomb_exchange

    eri = water["eri"]
    jk = exact_coulomb_exchange(eri)
    j_of, k_of = scf_ref.exact_operators(eri)
    rng = np.random.default_rng(7)
    a = rng.standard_normal((7, 7))
    z = rng.standard_normal((7, 7)) + 1j * rng.standard_normal((7, 7))
    for d in (a + a.T, a, z + z.conj().T):
        np.testing.assert_allclose(jk.coulomb(d), j_of(d), atol=STRUCTURE_TOL)
        np.testing.assert_allclose(jk.exchange(d), k_of(d), atol=STRUCTURE_TOL)
    with pytest.raises(InputError, match="ao, ao, ao, ao"):
        exact_coulomb_exchange(np.zeros((7, 7, 7)))


def test_exact_rhf_energy_of_water_matches_the_stored_value(water):
    """The GPU RHF driver on the exact builder reproduces the reference's stored (PySCF) exact
    RHF energy and orbital energies of water in STO-3G."""
    from hafnium_amd.density_fitting import exact_coulomb_exchange
    from hafnium_amd.scf import RestrictedHamiltonian, solve_scf

    e_nuc = scf_ref.nuclear_repulsion(water["charges"], water["positions_bohr"])
    h = RestrictedHamiltonian(water["overlap"], water["kinetic"] + water["nuclear_attraction"],
                              exact_coulomb_exchange(water["eri"]), 5, e_nuc)
    result = solve_scf(h)
    assert result.converged
    assert result.energy == pytest.approx(float(water["rhf_energy"]), abs=ENERGY_TOL)
    mo = result.mo_energies
    mo = mo.cpu().numpy() if hasattr(mo, "cpu") else np.asarray(mo)
    np.testing.assert_allclose(mo, water["rhf_mo_energies"], atol=MO_ENERGY_TOL)


@pytest.mark.parametrize("z", [8, 1])
def test_spherical_atom_matches_the_oracle(z):
    """The atoms of the density guess: spin-averaged restricted atoms with exact integrals
    (atoms.py:55-73) against the oracle's SCF with averaged occupation on the same integrals."""
    from hafnium_amd import integrals, molecule
    from hafnium_amd.clusters import atomic_density_guess

    atom = molecule.Molecule((z,), np.zeros((1, 3)))
    orb = molecule.BasisOnAtoms("cc-pvdz", atom)
    s, t, v = integrals.one_electron(orb, atom)
    j_of, k_of = scf_ref.exact_operators(integrals.four_center(orb))
    ref = scf_ref.rhf(s, t + v, j_of, k_of, z / 2.0, 0.0, occupation="averaged")
    assert ref.converged
    cache: dict = {}
    d = atomic_density_guess(atom, cache=cache)
    assert set(cache) == {z}
    np.testing.assert_allclose(d, d.T, atol=1e-14)
    assert float(np.trace(d @ s)) == pytest.approx(z / 2.0, abs=1e-10)
    np.testing.assert_allclose(d, ref.density, atol=1e-6)
    energy = 2.0 * float(np.sum(d * (t + v))) + float(np.sum(d * j_of(2.0 * d))) - float(np.sum(d * k_of(d)))
    assert energy == pytest.approx(ref.energy, abs=1e-9)


def test_water_guess_is_the_block_diagonal_superposition():
    from hafnium_amd.clusters import atomic_density_guess
    from hafnium_amd.molecule import water_cluster

    mol = water_cluster(1, 3)
    cache: dict = {}
    d = atomic_density_guess(mol, cache=cache)
    o, h = cache[8], cache[1]
    n_o, n_h = o.shape[0], h.shape[0]
    assert d.shape == (n_o + 2 * n_h,) * 2
    np.testing.assert_array_equal(d[:n_o, :n_o], o)
    np.testing.assert_array_equal(d[n_o:n_o + n_h, n_o:n_o + n_h], h)
    assert np.count_nonzero(d[:n_o, n_o:]) == 0
