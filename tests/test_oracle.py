"""The oracle pinned against the reference's stored fixtures, closed forms and brute force.

Each test is chosen so that a plausible mistake in the restatement (a dropped
term, a transposed operand, a wrong spin factor) fails it; none re-types the
oracle's own formula.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import jk_ref, layout_ref, scf_ref
from hafnium_amd import synthetic
from tests.tolerances import (
    DENSITY_FITTING_TOL,
    ENERGY_TOL,
    FITTED_INTEGRAL_TOL,
    MO_ENERGY_TOL,
    STRUCTURE_TOL,
)


def water_problem(water):
    overlap = water["overlap"]
    core = water["kinetic"] + water["nuclear_attraction"]
    e_nuc = scf_ref.nuclear_repulsion(water["charges"], water["positions_bohr"])
    return overlap, core, e_nuc


def test_nuclear_repulsion_matches_the_stored_geometry(water):
    # E_nuc of the stored geometry (Z = 8, 1, 1 at the stored Bohr positions), the value with which
    # the stored integrals reproduce the stored DF-RHF energy.
    _, _, e_nuc = water_problem(water)
    assert e_nuc == pytest.approx(9.189533762639684, abs=1e-12)


@pytest.mark.parametrize("method", ["eigh", "cholesky"])
def test_df_rhf_energy_matches_the_stored_pyscf_value(water, method):
    """PySCF DF-RHF water/STO-3G/def2-universal-jkfit (mp2_water_sto-3g.json)."""
    overlap, core, e_nuc = water_problem(water)
    b = scf_ref.df_tensor(water["metric"], water["three_center"], method)
    j_of, k_of = scf_ref.fitted_operators(b)
    result = scf_ref.rhf(overlap, core, j_of, k_of, 5, e_nuc)
    assert result.converged
    assert result.energy == pytest.approx(float(water["df_rhf_energy"]), abs=ENERGY_TOL)


def test_exact_rhf_energy_and_orbitals_match_the_stored_values(water):
    """scf_water_sto-3g.json from the stored four-centre integrals."""
    overlap, core, e_nuc = water_problem(water)
    j_of, k_of = scf_ref.exact_operators(water["eri"])
    result = scf_ref.rhf(overlap, core, j_of, k_of, 5, e_nuc)
    assert result.energy == pytest.approx(float(water["rhf_energy"]), abs=ENERGY_TOL)
    np.testing.assert_allclose(result.mo_energies, water["rhf_mo_energies"], atol=MO_ENERGY_TOL)
    # Converged state: symmetric, idempotent D S D = D, tr(D S) = n_alpha.
    d = result.density
    np.testing.assert_allclose(d, d.T, atol=STRUCTURE_TOL)
    np.testing.assert_allclose(d @ overlap @ d, d, atol=STRUCTURE_TOL)
    assert float(np.trace(d @ overlap)) == pytest.approx(5.0, abs=STRUCTURE_TOL)


def test_density_form_matches_the_looped_four_index_definitions(water):
    b = scf_ref.df_tensor(water["metric"], water["three_center"])
    eri = np.einsum("Puv,Pwx->uvwx", b, b)
    rng = np.random.default_rng(1)
    a = rng.standard_normal((7, 7))
    density = a + a.T
    j_loop, k_loop = jk_ref.looped_coulomb_exchange(eri, density)
    np.testing.assert_allclose(jk_ref.coulomb(b, density), j_loop, atol=STRUCTURE_TOL)
    np.testing.assert_allclose(jk_ref.exchange(b, density), k_loop, atol=STRUCTURE_TOL)
    # A non-symmetric density separates K[u,v] = sum (uw|xv) D[w,x] from its transpose.
    nonsym = rng.standard_normal((7, 7))
    _, k_ns = jk_ref.looped_coulomb_exchange(eri, nonsym)
    np.testing.assert_allclose(jk_ref.exchange(b, nonsym), k_ns, atol=STRUCTURE_TOL)
    # A complex Hermitian density (tests/unit/test_coulomb_exchange.py:45-51 of the reference):
    # J of it is real, K is Hermitian.
    z = rng.standard_normal((7, 7)) + 1j * rng.standard_normal((7, 7))
    herm = z + z.conj().T
    j_z, k_z = jk_ref.looped_coulomb_exchange(eri, herm)
    np.testing.assert_allclose(jk_ref.coulomb(b, herm), j_z, atol=STRUCTURE_TOL)
    np.testing.assert_allclose(jk_ref.exchange(b, herm), k_z, atol=STRUCTURE_TOL)
    np.testing.assert_allclose(np.imag(j_z), 0.0, atol=STRUCTURE_TOL)
    np.testing.assert_allclose(k_z, k_z.conj().T, atol=STRUCTURE_TOL)


def test_fitted_contractions_approximate_the_exact_ones(water):
    """tests/unit/test_coulomb_exchange.py:54-62 of the reference, real densities."""
    b = scf_ref.df_tensor(water["metric"], water["three_center"])
    assert b.shape[0] == 113
    j_of, k_of = scf_ref.exact_operators(water["eri"])
    rng = np.random.default_rng(2)
    a = rng.standard_normal((7, 7))
    density = a + a.T
    np.testing.assert_allclose(jk_ref.coulomb(b, density), j_of(density), atol=FITTED_INTEGRAL_TOL)
    np.testing.assert_allclose(jk_ref.exchange(b, density), k_of(density), atol=FITTED_INTEGRAL_TOL)


def test_fitted_integrals_equal_the_explicit_metric_solve(water):
    """sum_P B B = (uv|P) M^-1 (Q|wx) (tests/unit/test_density_fitting.py:177-195)."""
    metric, three = water["metric"], water["three_center"]
    flat = three.reshape(113, -1)
    explicit = flat.T @ np.linalg.solve(metric, flat)
    for method in ("eigh", "cholesky"):
        b = scf_ref.df_tensor(metric, three, method).reshape(113, -1)
        np.testing.assert_allclose(b.T @ b, explicit, atol=DENSITY_FITTING_TOL)
    b = scf_ref.df_tensor(metric, three)
    np.testing.assert_allclose(b, np.transpose(b, (0, 2, 1)), atol=DENSITY_FITTING_TOL)


def test_occupied_form_equals_density_form():
    b, c, d = synthetic.jk_inputs(3, 40, 37, 6)
    np.testing.assert_allclose(jk_ref.exchange_occupied(b, c), jk_ref.exchange(b, d), rtol=0, atol=1e-12)
    # gamma_Q = <C, B_Q C>
    x = jk_ref.half_transform(b, c)
    np.testing.assert_allclose(np.einsum("qmi,mi->q", x, c), jk_ref.charges(b, d), atol=1e-12)


def test_jk_invariants_for_a_density_of_orbitals():
    b, c, d = synthetic.jk_inputs(4, 30, 25, 5)
    j = jk_ref.coulomb(b, d)
    k = jk_ref.exchange(b, d)
    np.testing.assert_allclose(j, j.T, atol=1e-12)
    np.testing.assert_allclose(k, k.T, atol=1e-12)
    assert np.min(np.linalg.eigvalsh(k)) > -1e-10  # K is PSD for a PSD density
    e_j, e_k = jk_ref.energies(d, j, k)
    assert e_j >= 0.0 and e_k >= 0.0
    # Linearity of J in D.
    d2 = np.eye(25) * 0.3
    np.testing.assert_allclose(jk_ref.coulomb(b, d + d2), j + jk_ref.coulomb(b, d2), atol=1e-11)


def test_diis_solves_a_linear_model_exactly():
    """tests/unit/test_scf.py:501-538 of the reference."""
    focks = [np.array([[x, 0.0], [0.0, 1.0]]) for x in (1.0, 3.0)]
    errors = [np.array([[x - 2.0, 0.0], [0.0, 0.0]]) for x in (1.0, 3.0)]
    diis = scf_ref.Diis(2)
    diis.push(focks[0], errors[0])
    assert diis.extrapolate() is focks[0]
    diis.push(focks[1], errors[1])
    np.testing.assert_allclose(diis.extrapolate(), [[2.0, 0.0], [0.0, 1.0]], atol=STRUCTURE_TOL)
    dependent = scf_ref.Diis(3)
    for f, e in ((focks[0], errors[0]), (focks[1], errors[1]), (focks[1], errors[1])):
        dependent.push(f, e)
    np.testing.assert_allclose(dependent.extrapolate(), [[2.0, 0.0], [0.0, 1.0]], atol=STRUCTURE_TOL)
    zero = np.zeros((2, 2))
    settled = scf_ref.Diis(2)
    settled.push(focks[0], zero)
    settled.push(focks[1], zero)
    np.testing.assert_allclose(settled.extrapolate(), [[2.0, 0.0], [0.0, 1.0]], atol=STRUCTURE_TOL)


def test_canonical_orthogonalizer_orthonormalises(water):
    x = scf_ref.canonical_orthogonalizer(water["overlap"])
    np.testing.assert_allclose(x.T @ water["overlap"] @ x, np.eye(x.shape[1]), atol=STRUCTURE_TOL)
    singular = np.ones((3, 3))
    x = scf_ref.canonical_orthogonalizer(singular)
    assert x.shape == (3, 1)


def test_degenerate_groups_share_the_frontier():
    occ = scf_ref.fill_degenerate_groups(np.array([-1.0, 0.5, 0.5, 2.0]), 2.0)
    np.testing.assert_allclose(occ, [1.0, 0.5, 0.5, 0.0])
    assert list(scf_ref.fill_whole_levels(np.array([0.3, -1.0, 2.0]), 2)) == [1.0, 1.0, 0.0]


def test_a_planted_mistake_in_the_oracle_fails_the_energy_check(water):
    """The stored DF energy catches a dropped spin factor in J (2 D_a -> D_a)."""
    overlap, core, e_nuc = water_problem(water)
    b = scf_ref.df_tensor(water["metric"], water["three_center"])
    wrong_j = lambda d: jk_ref.coulomb(b, 0.5 * d)  # noqa: E731
    _, k_of = scf_ref.fitted_operators(b)
    result = scf_ref.rhf(overlap, core, wrong_j, k_of, 5, e_nuc)
    assert abs(result.energy - float(water["df_rhf_energy"])) > 1e-3


def test_tiled_layout_round_trip_and_weights():
    rng = np.random.default_rng(7)
    for n in (1, 31, 32, 33, 70):
        a = rng.standard_normal((3, n, n))
        b = a + np.transpose(a, (0, 2, 1))
        packed = layout_ref.pack(b)
        assert packed.shape == (3, layout_ref.packed_elems(n))
        np.testing.assert_array_equal(layout_ref.unpack(packed, n), b)
        d = rng.standard_normal((n, n))  # need not be symmetric
        w = layout_ref.weights(d)
        np.testing.assert_allclose(packed @ w, jk_ref.charges(b, d), atol=1e-11)


def test_averaged_occupation_gives_a_spherical_oxygen_atom():
    """The oracle's spherical atom (the reference's _spherical_atom, atoms.py:55-73): exact
    integrals of O in cc-pVDZ, four alpha electrons averaged over the degenerate 2p level. Its
    occupations are (1, 1, 2/3, 2/3, 2/3), tr(D S) = 4, and every block of D between two p shells
    is a multiple of the identity (a spherical density); whole-level filling breaks that."""
    from hafnium_amd import integrals, molecule

    atom = molecule.Molecule((8,), np.zeros((1, 3)))
    orb = molecule.BasisOnAtoms("cc-pvdz", atom)
    s, t, v = integrals.one_electron(orb, atom)
    j_of, k_of = scf_ref.exact_operators(integrals.four_center(orb))
    res = scf_ref.rhf(s, t + v, j_of, k_of, 4.0, 0.0, occupation="averaged")
    assert res.converged
    np.testing.assert_allclose(np.sort(res.occupations)[::-1][:6], [1, 1, 2 / 3, 2 / 3, 2 / 3, 0], atol=1e-12)
    assert float(np.trace(res.density @ s)) == pytest.approx(4.0, abs=1e-10)
    p_shells, at = [], 0
    for l, count in zip(orb._l, orb._ncontr):
        for _ in range(count):
            if l == 1:
                p_shells.append(list(range(at, at + 3)))
            at += 2 * l + 1

    def spherical(d):
        return all(np.allclose(d[np.ix_(a, b)], d[a[0], b[0]] * np.eye(3), atol=1e-8)
                   for a in p_shells for b in p_shells)

    assert spherical(res.density)
    whole = scf_ref.rhf(s, t + v, j_of, k_of, 4, 0.0)
    assert not spherical(whole.density)


def test_exact_uhf_of_the_cation_matches_the_stored_values(water):
    """The stored PySCF UHF of the water cation (9 electrons, doublet) from the stored ERIs."""
    overlap, core, e_nuc = water_problem(water)
    j_of, k_of = scf_ref.exact_operators(water["eri"])
    result = scf_ref.uhf(overlap, core, j_of, k_of, 5, 4, e_nuc, energy_tolerance=1e-12,
                         gradient_tolerance=1e-8)
    assert result.converged
    assert result.energy == pytest.approx(float(water["cation_uhf_energy"]), abs=ENERGY_TOL)
    for s in range(2):
        np.testing.assert_allclose(result.mo_energies[s], water["cation_uhf_mo_energies"][s], atol=MO_ENERGY_TOL)
        np.testing.assert_array_equal(result.occupations[s], water["cation_uhf_mo_occupations"][s])
    assert result.spin_square == pytest.approx(float(water["cation_uhf_spin_square"]), abs=1e-8)
    # The same-spin exchange only: K of the total density on both blocks misses the stored energy.
    wrong = scf_ref.uhf(overlap, core, j_of, lambda d: 2.0 * k_of(d), 5, 4, e_nuc)
    assert abs(wrong.energy - float(water["cation_uhf_energy"])) > 1e-2
