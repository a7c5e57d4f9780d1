"""DF-RHF of water clusters on the GPU: inputs, the atomic-density guess, timed runs.

The workload of the SCF time-to-solution measurements: a seeded (H2O)_n
(``molecule.water_cluster``), cc-pVDZ with the def2-universal-jkfit
auxiliary basis (cc-pVDZ-JKFIT is not bundled offline), integrals from the
library's engine, the fitted tensor by the GPU Cholesky route, and
``solve_scf`` in plain FP64 or with the mixed-precision policy.

``atomic_density_guess`` mirrors the reference's superposition of atomic
densities (hafnium/scf/atoms.py:55-119): one spin-averaged spherical atomic
RHF per element (``AveragedOccupation`` with half the electrons per spin),
the atomic blocks placed on the diagonal, each atom converged with exact
four-centre integrals as in the reference.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from hafnium_amd import integrals
from hafnium_amd.molecule import BasisOnAtoms, Molecule, water_cluster
from hafnium_amd.scf import RestrictedHamiltonian, ScfOptions, solve_scf

ORBITAL_BASIS = "cc-pvdz"
AUX_BASIS = "def2-universal-jkfit"


def _torch():
    import torch

    return torch


@dataclass
class ClusterSystem:
    molecule: Molecule
    hamiltonian: RestrictedHamiltonian
    timings: dict = field(default_factory=dict)
    metric: np.ndarray | None = None         # (P|Q) [naux, naux], host copy (keep_integrals)
    three_center: np.ndarray | None = None   # (P|uv) [naux, n, n], host copy (keep_integrals)
    aux_basis: str = AUX_BASIS


def build_system(molecule: Molecule, *, basis: str = ORBITAL_BASIS, aux: str = AUX_BASIS,
                 precision: str = "fp64", device=None, keep_integrals: bool = False,
                 precomputed=None, device_integrals: bool = True) -> ClusterSystem:
    """Integrals, fitted tensor (GPU) and the Hamiltonian of a closed shell.

    The metric and the three-centre integrals are computed on the GPU
    (``device_integrals``, the default) or by the host engine; S, T and V on
    the host. ``keep_integrals`` keeps host copies of the metric and the
    three-centre integrals on the result (the learned guess needs them).
    ``precomputed`` = (S, T + V, metric, three-centre) skips the integral
    step."""
    from hafnium_amd.density_fitting import fitted_coulomb_exchange

    torch = _torch()
    t0 = time.perf_counter()
    if precomputed is None:
        orb = BasisOnAtoms(basis, molecule)
        auxb = BasisOnAtoms(aux, molecule)
        s, t, v = integrals.one_electron(orb, molecule)
        core = t + v
        if device_integrals:
            m = integrals.metric_device(auxb, device=device)
            tc = integrals.three_center_device(auxb, orb, device=device)
            torch.cuda.synchronize()
        else:
            m = integrals.metric(auxb)
            tc = integrals.three_center(auxb, orb)
    else:
        s, core, m, tc = precomputed
    t1 = time.perf_counter()
    two_body = fitted_coulomb_exchange(m, tc, precision=precision, device=device)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    if keep_integrals and isinstance(m, torch.Tensor):
        m, tc = m.cpu().numpy(), tc.cpu().numpy()
    if molecule.n_electrons % 2:
        n_alpha = molecule.n_electrons / 2.0
    else:
        n_alpha = molecule.n_electrons // 2
    h = RestrictedHamiltonian(s, core, two_body, n_alpha, molecule.nuclear_repulsion())
    return ClusterSystem(molecule, h, {"integrals_s": t1 - t0, "fit_s": t2 - t1,
                                       "nao": s.shape[0], "naux": m.shape[0]},
                         m if keep_integrals else None, tc if keep_integrals else None, aux)


def atomic_density_guess(molecule: Molecule, *, basis: str = ORBITAL_BASIS, device=None,
                         cache: dict | None = None):
    """Block-diagonal superposition of spin-averaged atomic densities (D_alpha).

    As the reference's ``atomic_density_guess`` (hafnium/scf/atoms.py:76-119, ``_spherical_atom``
    :55-73): every element is converged once as the neutral atom at the origin in its own shells,
    restricted, with its electrons averaged over both spins and over degenerate levels
    (occupation "averaged"), with exact four-centre integrals and the default SCF options; the
    atomic densities go on the diagonal in the order of the atoms. ``cache`` maps the atomic
    number to the atom's density (one basis per cache).
    """
    from hafnium_amd.density_fitting import exact_coulomb_exchange

    cache = {} if cache is None else cache
    blocks = []
    for z in molecule.charges:
        if z not in cache:
            atom = Molecule((z,), np.zeros((1, 3)))
            orb = BasisOnAtoms(basis, atom)
            s, t, v = integrals.one_electron(orb, atom)
            two_body = exact_coulomb_exchange(integrals.four_center(orb), device=device)
            h = RestrictedHamiltonian(s, t + v, two_body, z / 2.0, 0.0)
            result = solve_scf(h, occupation="averaged", options=ScfOptions())
            cache[z] = result.density.cpu().numpy()
        blocks.append(cache[z])
    n = sum(b.shape[0] for b in blocks)
    out = np.zeros((n, n))
    at = 0
    for b in blocks:
        out[at:at + b.shape[0], at:at + b.shape[0]] = b
        at += b.shape[0]
    return out


def run_cluster(n_waters: int, seed: int, *, precision: str = "fp64", guess=None, energy_tolerance: float = 1e-8,
                gradient_tolerance: float = 1e-6, switch_threshold: float = 1e-5, device=None, system=None):
    """Converge (H2O)_n; returns ``(result, system)`` with the SCF wall time in ``result.seconds``.

    ``guess``: a density, or a start of ``learned_guess.learned_guess`` (density, orbitals and the
    first Fock build)."""
    if system is None:
        system = build_system(water_cluster(n_waters, seed), precision=precision, device=device)
    opts = ScfOptions(precision=precision, energy_tolerance=energy_tolerance, gradient_tolerance=gradient_tolerance,
                      switch_threshold=switch_threshold)
    _torch().cuda.synchronize()
    if hasattr(guess, "first_build"):
        result = solve_scf(system.hamiltonian, guess=guess.density, guess_orbitals=guess.orbitals,
                           first_build=guess.first_build, options=opts)
    else:
        result = solve_scf(system.hamiltonian, guess=guess, options=opts)
    return result, system
