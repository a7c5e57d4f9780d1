"""The GPU RHF driver and the GPU fitted tensor against the reference's stored values."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import scf_ref
from tests.tolerances import DENSITY_FITTING_TOL, ENERGY_TOL, STRUCTURE_TOL

pytestmark = pytest.mark.gpu


def water_hamiltonian(water, precision="fp64"):
    from hafnium_amd.density_fitting import fitted_coulomb_exchange
    from hafnium_amd.scf import RestrictedHamiltonian

    two_body = fitted_coulomb_exchange(water["metric"], water["three_center"], precision=precision)
    e_nuc = scf_ref.nuclear_repulsion(water["charges"], water["positions_bohr"])
    return RestrictedHamiltonian(water["overlap"], water["kinetic"] + water["nuclear_attraction"], two_body, 5, e_nuc)


def test_gpu_cholesky_fit_reproduces_the_metric_solve(water):
    import torch

    from hafnium_amd.coulomb_exchange import _Context
    from hafnium_amd.density_fitting import df_tensor
    from hafnium_amd.linalg import Dense

    dev = torch.device("cuda", 0)
    b = df_tensor(Dense(_Context(0), dev), water["metric"], water["three_center"]).cpu().numpy()
    flat = water["three_center"].reshape(113, -1)
    explicit = flat.T @ np.linalg.solve(water["metric"], flat)
    np.testing.assert_allclose(b.reshape(113, -1).T @ b.reshape(113, -1), explicit, atol=DENSITY_FITTING_TOL)
    np.testing.assert_allclose(b, scf_ref.df_tensor(water["metric"], water["three_center"], "cholesky"), atol=1e-12)


def test_dense_toolbox_matches_numpy():
    import torch

    from hafnium_amd.coulomb_exchange import _Context
    from hafnium_amd.linalg import Dense

    dev = torch.device("cuda", 0)
    dn = Dense(_Context(0), dev)
    rng = np.random.default_rng(3)
    a = rng.standard_normal((37, 21))
    b = rng.standard_normal((21, 50))
    c = rng.standard_normal((50, 21))
    ta, tb, tc = (torch.from_numpy(x).to(dev) for x in (a, b, c))
    np.testing.assert_allclose(dn.gemm(ta, tb).cpu().numpy(), a @ b, atol=1e-12)
    np.testing.assert_allclose(dn.gemm(ta, tc, tb=True).cpu().numpy(), a @ c.T, atol=1e-12)
    np.testing.assert_allclose(dn.gemm(tb, tc, ta=True, tb=True).cpu().numpy(), b.T @ c.T, atol=1e-12)
    s = a.T @ a
    w, v = dn.eigh(torch.from_numpy(s).to(dev))
    np.testing.assert_allclose(w.cpu().numpy(), np.linalg.eigvalsh(s), rtol=1e-12)
    vh = v.cpu().numpy()
    np.testing.assert_allclose(vh.T @ s @ vh, np.diag(w.cpu().numpy()), atol=1e-9)
    out = torch.empty(1, dtype=torch.float64, device=dev)
    dn.dot(ta, ta, out)
    assert out.item() == pytest.approx(float(np.sum(a * a)), rel=1e-14)
    m = torch.from_numpy(rng.standard_normal((9, 9))).to(dev)
    np.testing.assert_allclose(dn.antisymmetrize(m).cpu().numpy(), (m - m.T).cpu().numpy(), atol=0)


@pytest.mark.parametrize("precision", ["fp64", "bf16x3", "fp32"])
def test_df_rhf_water_matches_the_stored_energy(water, precision):
    from hafnium_amd.scf import ScfOptions, solve_scf

    h = water_hamiltonian(water)
    result = solve_scf(h, options=ScfOptions(precision=precision))
    assert result.converged
    assert result.energy == pytest.approx(float(water["df_rhf_energy"]), abs=ENERGY_TOL)
    assert result.history[-1].precision == "fp64"
    d = result.density.cpu().numpy()
    s = water["overlap"]
    np.testing.assert_allclose(d, d.T, atol=STRUCTURE_TOL)
    np.testing.assert_allclose(d @ s @ d, d, atol=STRUCTURE_TOL)
    assert float(np.trace(d @ s)) == pytest.approx(5.0, abs=STRUCTURE_TOL)


def test_non_convergence_raises_with_the_last_state(water):
    from hafnium_amd.errors import ConvergenceError
    from hafnium_amd.scf import ScfOptions, solve_scf

    h = water_hamiltonian(water)
    with pytest.raises(ConvergenceError) as info:
        solve_scf(h, options=ScfOptions(max_iterations=2))
    last = info.value.last
    assert not last.converged and last.n_iterations == 2
    restarted = solve_scf(h, guess=last.density)
    assert restarted.energy == pytest.approx(float(water["df_rhf_energy"]), abs=ENERGY_TOL)


def duplicated_aux(water, k=17):
    """The stored metric and three-centre integrals with aux function ``k`` appended a second time:
    the fitted space is unchanged, the metric exactly singular."""
    m, t = water["metric"], water["three_center"]
    n = m.shape[0]
    m2 = np.zeros((n + 1, n + 1))
    m2[:n, :n] = m
    m2[n, :n] = m[k]
    m2[:n, n] = m[:, k]
    m2[n, n] = m[k, k]
    return m2, np.concatenate([t, t[k:k + 1]], axis=0)


def test_singular_metric_drops_eigenvalues_like_the_reference(water):
    """The eigenvalue-dropping fit (density_fitting.py:181-197): a duplicated aux function leaves
    the fitted space, J, K and the DF-RHF energy unchanged."""
    from hafnium_amd.density_fitting import fitted_coulomb_exchange
    from hafnium_amd.scf import RestrictedHamiltonian, solve_scf
    from oracle import jk_ref

    m2, t2 = duplicated_aux(water)
    jk = fitted_coulomb_exchange(m2, t2)
    assert jk.n_aux == 113  # the null direction is dropped
    b_ref = scf_ref.df_tensor(m2, t2, "eigh")
    rng = np.random.default_rng(8)
    d = rng.standard_normal((7, 7))
    d = d + d.T
    assert jk_ref.max_relative_error(jk.coulomb(d), jk_ref.coulomb(b_ref, d)) < 1e-10
    assert jk_ref.max_relative_error(jk.exchange(d), jk_ref.exchange(b_ref, d)) < 1e-10
    e_nuc = scf_ref.nuclear_repulsion(water["charges"], water["positions_bohr"])
    ham = RestrictedHamiltonian(water["overlap"], water["kinetic"] + water["nuclear_attraction"], jk, 5, e_nuc)
    assert solve_scf(ham).energy == pytest.approx(float(water["df_rhf_energy"]), abs=ENERGY_TOL)


@pytest.mark.parametrize("method", ["auto", "cholesky", "eigh"])
def test_every_fitting_method_gives_the_stored_df_energy(water, method):
    from hafnium_amd.density_fitting import fitted_coulomb_exchange
    from hafnium_amd.scf import RestrictedHamiltonian, solve_scf

    jk = fitted_coulomb_exchange(water["metric"], water["three_center"], method=method)
    assert jk.n_aux == 113
    e_nuc = scf_ref.nuclear_repulsion(water["charges"], water["positions_bohr"])
    ham = RestrictedHamiltonian(water["overlap"], water["kinetic"] + water["nuclear_attraction"], jk, 5, e_nuc)
    assert solve_scf(ham).energy == pytest.approx(float(water["df_rhf_energy"]), abs=ENERGY_TOL)
