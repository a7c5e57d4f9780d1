"""Spin structures through the GPU kernels: the collinear and spinor assembly of J and K, and the
unrestricted SCF of the water cation against the stored PySCF values."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import jk_ref, scf_ref
from tests.tolerances import ENERGY_TOL, MO_ENERGY_TOL, STRUCTURE_TOL

pytestmark = pytest.mark.gpu


def random_hermitian(rng, n):
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return 0.5 * (a + a.conj().T)


@pytest.fixture(scope="module")
def exact(water):
    from hafnium_amd.density_fitting import exact_coulomb_exchange

    return exact_coulomb_exchange(water["eri"])


def looped(water, d):
    return jk_ref.looped_coulomb_exchange(water["eri"], d)


def test_collinear_structure_uses_the_total_density_for_coulomb(water, exact):
    from hafnium_amd.coulomb_exchange import SpinDensity, coulomb_exchange_matrices

    rng = np.random.default_rng(3)
    alpha, beta = random_hermitian(rng, 7), random_hermitian(rng, 7)
    j, k = coulomb_exchange_matrices(exact, SpinDensity("unrestricted", (alpha, beta)))
    total, _ = looped(water, np.real(alpha + beta))
    for block in j.blocks:
        np.testing.assert_allclose(block, total, atol=STRUCTURE_TOL)
    np.testing.assert_allclose(k.blocks[0], looped(water, alpha)[1], atol=STRUCTURE_TOL)
    np.testing.assert_allclose(k.blocks[1], looped(water, beta)[1], atol=STRUCTURE_TOL)


def test_restricted_structure_doubles_the_stored_block(water, exact):
    from hafnium_amd.coulomb_exchange import SpinDensity, coulomb_exchange_matrices

    alpha = np.real(random_hermitian(np.random.default_rng(4), 7))
    j, k = coulomb_exchange_matrices(exact, SpinDensity("restricted", (alpha,)))
    np.testing.assert_allclose(j.blocks[0], looped(water, 2 * alpha)[0], atol=STRUCTURE_TOL)
    np.testing.assert_allclose(k.blocks[0], looped(water, alpha)[1], atol=STRUCTURE_TOL)


def test_spinor_structure_contracts_every_spin_block(water, exact):
    from hafnium_amd.coulomb_exchange import SpinDensity, coulomb_exchange_matrices

    n = 7
    spinor = random_hermitian(np.random.default_rng(5), 2 * n)
    j, k = coulomb_exchange_matrices(exact, SpinDensity("generalized", (spinor,)))
    total, _ = looped(water, np.real(spinor[:n, :n] + spinor[n:, n:]))
    np.testing.assert_allclose(j.blocks[0], np.kron(np.eye(2), total), atol=STRUCTURE_TOL)
    expected = np.block([[looped(water, spinor[:n, :n])[1], looped(water, spinor[:n, n:])[1]],
                         [looped(water, spinor[n:, :n])[1], looped(water, spinor[n:, n:])[1]]])
    np.testing.assert_allclose(k.blocks[0], expected, atol=STRUCTURE_TOL)
    np.testing.assert_allclose(k.blocks[0], k.blocks[0].conj().T, atol=STRUCTURE_TOL)


def test_spinor_structure_with_device_tensors_and_the_fitted_builder(water):
    """Torch inputs stay torch; the fitted builder's spinor blocks match the oracle's density form."""
    import torch

    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange, SpinDensity, coulomb_exchange_matrices

    b = scf_ref.df_tensor(water["metric"], water["three_center"], "cholesky")
    jk = MI355XCoulombExchange(b)
    n = 7
    spinor = random_hermitian(np.random.default_rng(6), 2 * n)
    j, k = coulomb_exchange_matrices(jk, SpinDensity("generalized", (torch.from_numpy(spinor).cuda(),)))
    assert isinstance(k.blocks[0], torch.Tensor)
    kh = k.blocks[0].cpu().numpy()
    for s in range(2):
        for t in range(2):
            block = spinor[s * n:(s + 1) * n, t * n:(t + 1) * n]
            expect = jk_ref.exchange(b, np.real(block)) + 1j * jk_ref.exchange(b, np.imag(block))
            np.testing.assert_allclose(kh[s * n:(s + 1) * n, t * n:(t + 1) * n], expect, atol=STRUCTURE_TOL)
    expect_j = jk_ref.coulomb(b, np.real(spinor[:n, :n] + spinor[n:, n:]))
    np.testing.assert_allclose(j.blocks[0].cpu().numpy(), np.kron(np.eye(2), expect_j), atol=STRUCTURE_TOL)


def water_hamiltonian(water, two_body, n_alpha, n_beta):
    from hafnium_amd.uhf import UnrestrictedHamiltonian

    e_nuc = scf_ref.nuclear_repulsion(water["charges"], water["positions_bohr"])
    return UnrestrictedHamiltonian(water["overlap"], water["kinetic"] + water["nuclear_attraction"], two_body,
                                   n_alpha, n_beta, e_nuc)


def test_cation_uhf_matches_the_stored_pyscf_values(water, exact):
    from hafnium_amd.scf import ScfOptions
    from hafnium_amd.uhf import solve_uhf

    result = solve_uhf(water_hamiltonian(water, exact, 5, 4),
                       options=ScfOptions(energy_tolerance=1e-12, gradient_tolerance=1e-8))
    assert result.converged
    assert result.energy == pytest.approx(float(water["cation_uhf_energy"]), abs=ENERGY_TOL)
    for s in range(2):
        np.testing.assert_allclose(result.mo_energies[s].cpu().numpy(), water["cation_uhf_mo_energies"][s],
                                   atol=MO_ENERGY_TOL)
        np.testing.assert_array_equal(result.mo_occupations[s], water["cation_uhf_mo_occupations"][s])
    assert result.spin_square == pytest.approx(float(water["cation_uhf_spin_square"]), abs=1e-8)


def test_fitted_cation_uhf_matches_the_oracle(water):
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange
    from hafnium_amd.uhf import solve_uhf

    b = scf_ref.df_tensor(water["metric"], water["three_center"], "cholesky")
    ham = water_hamiltonian(water, MI355XCoulombExchange(b), 5, 4)
    got = solve_uhf(ham)
    j_of, k_of = scf_ref.fitted_operators(b)
    ref = scf_ref.uhf(water["overlap"], water["kinetic"] + water["nuclear_attraction"], j_of, k_of, 5, 4,
                      ham.nuclear_repulsion)
    assert ref.converged and got.converged
    assert got.energy == pytest.approx(ref.energy, abs=ENERGY_TOL)
    assert got.spin_square == pytest.approx(ref.spin_square, abs=1e-8)
    assert abs(got.n_iterations - len(ref.history)) <= 1


def test_closed_shell_uhf_is_the_rhf_state(water, exact):
    from hafnium_amd.uhf import solve_uhf

    result = solve_uhf(water_hamiltonian(water, exact, 5, 5))
    assert result.energy == pytest.approx(float(water["rhf_energy"]), abs=ENERGY_TOL)
    assert result.spin_square == pytest.approx(0.0, abs=1e-8)
    d_a, d_b = (d.cpu().numpy() for d in result.density)
    np.testing.assert_allclose(d_a, d_b, atol=1e-7)
