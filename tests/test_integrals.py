"""The integral engine against the reference's stored PySCF integrals (host code, CPU suite)."""

from __future__ import annotations

import os

import numpy as np
import pytest

from hafnium_amd import integrals, molecule
from oracle import scf_ref
from tests.conftest import GOLDEN
from tests.tolerances import ENERGY_TOL, FITTED_INTEGRAL_TOL

INTEGRAL_TOL = 1e-12  # the reference's tests/tolerances.py INTEGRAL_TOL


@pytest.fixture(scope="module")
def water_mol(water):
    return integrals.water_from_fixture_positions(water["positions_bohr"])


def test_sto3g_one_electron_integrals(water, water_mol):
    orb = molecule.BasisOnAtoms("sto-3g", water_mol)
    s, t, v = integrals.one_electron(orb, water_mol)
    np.testing.assert_allclose(s, water["overlap"], atol=INTEGRAL_TOL)
    np.testing.assert_allclose(t, water["kinetic"], atol=INTEGRAL_TOL)
    np.testing.assert_allclose(v, water["nuclear_attraction"], atol=INTEGRAL_TOL * 100)


def test_ccpvdz_one_electron_integrals():
    """cc-pVDZ: general contractions and d functions (24 functions)."""
    data = np.load(os.path.join(GOLDEN, "water_ccpvdz_one_electron.npz"))
    mol = integrals.water_from_fixture_positions(data["positions_bohr"])
    orb = molecule.BasisOnAtoms("cc-pvdz", mol)
    assert orb.n_functions == 24
    s, t, v = integrals.one_electron(orb, mol)
    np.testing.assert_allclose(s, data["overlap"], atol=INTEGRAL_TOL)
    np.testing.assert_allclose(t, data["kinetic"], atol=INTEGRAL_TOL * 10)
    np.testing.assert_allclose(v, data["nuclear_attraction"], atol=INTEGRAL_TOL * 100)


def test_density_fitting_integrals_match_the_stored_metric_and_three_centre(water, water_mol):
    """def2-universal-jkfit up to g functions: (P|Q) [113, 113] and (P|uv) [113, 7, 7]."""
    orb = molecule.BasisOnAtoms("sto-3g", water_mol)
    aux = molecule.BasisOnAtoms("def2-universal-jkfit", water_mol)
    assert aux.n_functions == 113
    m = integrals.metric(aux)
    tc = integrals.three_center(aux, orb)
    np.testing.assert_allclose(m, water["metric"], atol=INTEGRAL_TOL * 1000, rtol=1e-12)
    np.testing.assert_allclose(tc, water["three_center"], atol=INTEGRAL_TOL * 10)


def test_four_centre_integrals_match_the_stored_eri(water, water_mol):
    """(uv|wx) [7, 7, 7, 7] of water in STO-3G against the reference's stored PySCF values, and the
    eight-fold permutational symmetry."""
    orb = molecule.BasisOnAtoms("sto-3g", water_mol)
    eri = integrals.four_center(orb)
    np.testing.assert_allclose(eri, water["eri"], atol=INTEGRAL_TOL)
    for perm in ((1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)):
        np.testing.assert_array_equal(eri, eri.transpose(perm))


def test_four_centre_integrals_of_an_oxygen_atom():
    """cc-pVDZ oxygen (general contractions, d functions): the [n^2, n^2] matrix (uv|wx) is
    symmetric positive semidefinite, independent of where the atom sits, and approximated by the
    density fit of the (separately pinned) metric and three-centre integrals to the reference's
    fitted-integral tolerance (tests/unit/test_coulomb_exchange.py:54-62 makes the same comparison
    for water)."""
    out = []
    for position in (np.zeros((1, 3)), np.array([[0.3, -1.2, 2.5]])):
        atom = molecule.Molecule((8,), position)
        orb = molecule.BasisOnAtoms("cc-pvdz", atom)
        aux = molecule.BasisOnAtoms("def2-universal-jkfit", atom)
        n = orb.n_functions
        eri = integrals.four_center(orb)
        mat = eri.reshape(n * n, n * n)
        np.testing.assert_allclose(mat, mat.T, atol=1e-14)
        w = np.linalg.eigvalsh(mat)
        assert w.min() > -1e-12 * w.max()
        b = scf_ref.df_tensor(integrals.metric(aux), integrals.three_center(aux, orb), "cholesky")
        fitted = np.einsum("Puv,Pwx->uvwx", b, b)
        np.testing.assert_allclose(fitted, eri, atol=FITTED_INTEGRAL_TOL)
        out.append(eri)
    np.testing.assert_allclose(out[1], out[0], atol=1e-14)


def test_df_rhf_energy_from_own_integrals(water, water_mol):
    """The whole DF-RHF input chain reproduces the stored PySCF DF-RHF energy."""
    orb = molecule.BasisOnAtoms("sto-3g", water_mol)
    aux = molecule.BasisOnAtoms("def2-universal-jkfit", water_mol)
    s, t, v = integrals.one_electron(orb, water_mol)
    b = scf_ref.df_tensor(integrals.metric(aux), integrals.three_center(aux, orb), "cholesky")
    j_of, k_of = scf_ref.fitted_operators(b)
    result = scf_ref.rhf(s, t + v, j_of, k_of, 5, water_mol.nuclear_repulsion())
    assert result.energy == pytest.approx(float(water["df_rhf_energy"]), abs=ENERGY_TOL)


def test_integral_invariances():
    """Translation invariance, symmetry and positive-definiteness on a water dimer."""
    mol = molecule.water_cluster(2, seed=3)
    shifted = molecule.Molecule(mol.charges, mol.positions + np.array([1.3, -0.7, 2.1]))
    orb, orb2 = molecule.BasisOnAtoms("cc-pvdz", mol), molecule.BasisOnAtoms("cc-pvdz", shifted)
    s, t, v = integrals.one_electron(orb, mol)
    s2, t2, v2 = integrals.one_electron(orb2, shifted)
    np.testing.assert_allclose(s2, s, atol=1e-12)
    np.testing.assert_allclose(t2, t, atol=1e-11)
    np.testing.assert_allclose(v2, v, atol=1e-10)
    np.testing.assert_allclose(np.diag(s), 1.0, atol=1e-13)
    assert np.min(np.linalg.eigvalsh(s)) > 0
    aux = molecule.BasisOnAtoms("def2-universal-jkfit", mol)
    m = integrals.metric(aux)
    np.testing.assert_allclose(m, m.T, atol=1e-13)
    assert np.min(np.linalg.eigvalsh(m)) > 0


def test_closed_form_s_functions():
    """(s|s) overlap and Coulomb of unit s Gaussians in closed form."""
    a, b, r = 0.8, 1.3, 1.7
    mol = molecule.Molecule((1, 1), np.array([[0.0, 0.0, 0.0], [0.0, 0.0, r]]))

    class OneS(molecule.BasisOnAtoms):
        def __init__(self, exps):
            self._l = np.zeros(2, dtype=np.int32)
            self._nprim = np.ones(2, dtype=np.int32)
            self._ncontr = np.ones(2, dtype=np.int32)
            self._centers = mol.positions.reshape(-1).copy()
            self._exps = np.asarray(exps, dtype=np.float64)
            self._coefs = np.ones(2)
            self.n_functions = 2

    basis = OneS([a, b])
    s, _, _ = integrals.one_electron(basis, molecule.Molecule((), np.zeros((0, 3))))
    p = a + b
    expected = (2 * np.sqrt(a * b) / p) ** 1.5 * np.exp(-a * b / p * r * r)
    assert s[0, 1] == pytest.approx(expected, rel=1e-13)
    m = integrals.metric(basis)
    # Coulomb of two normalised s Gaussians: (2/sqrt(pi)) (norms) * ... via erf.
    from math import erf, pi, sqrt

    na = (2 * a / pi) ** 0.75
    nb = (2 * b / pi) ** 0.75
    alpha = a * b / p
    coul = na * nb * (pi / a) ** 1.5 * (pi / b) ** 1.5 * erf(sqrt(alpha) * r) / r
    assert m[0, 1] == pytest.approx(coul, rel=1e-12)
