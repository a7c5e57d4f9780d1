"""Unrestricted (collinear) Hartree-Fock on the GPU, through the same J/K kernels.

The collinear structure of the reference's mean-field driver
(hafnium/scf/driver.py:154-213 with the unrestricted ``BlockStructure``):

* start: the spin-averaged core-Hamiltonian guess, half the electrons in
  each spin with the averaged filling (driver.py:216-246), so the first Fock
  build separates the spins;
* Fock: ``F_s = h + J[D_a + D_b] - K[D_s]`` (``coulomb_matrix`` /
  ``exchange_matrix``, coulomb_exchange.py:407-462: J of the spin trace on
  both blocks, K per block);
* energy: ``tr(D h) + tr(D J) / 2 - (tr(D_a K_a) + tr(D_b K_b)) / 2`` with
  ``D = D_a + D_b`` (hamiltonian.py:331-370, weight 1 per collinear block);
* gradient ``X^T (F_s D_s S - S D_s F_s) X`` per block, its norm and the
  DIIS overlaps summed over both blocks (driver.py:248-260, :385-395;
  diis.py:27-106), one set of DIIS coefficients for both Fock blocks;
* aufbau filling of ``n_alpha`` and ``n_beta`` levels (occupation.py:63-130);
* ``<S^2> = S_z (S_z + 1) + n_b - tr(D_a S D_b S)`` of the final state.

Every matrix stays on the device; one host transfer per iteration carries the
energy terms, the newest error's overlaps and the eigensolver status. The
exchange of an aufbau density runs in the occupied form on its orbitals; the
averaged first density goes through the density form (factorised on the
GPU). FP64 throughout.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from hafnium_amd.errors import BackendError, ConvergenceError, InputError
from hafnium_amd.linalg import Dense
from hafnium_amd.scf import EnergyTerms, ScfIteration, ScfOptions, _Setup, fill_degenerate_groups, solve_diis


def _torch():
    import torch

    return torch


def _tensor(x, device):
    torch = _torch()
    return torch.as_tensor(np.asarray(x) if isinstance(x, np.ndarray) else x,
                           dtype=torch.float64).to(device).contiguous()


class UnrestrictedHamiltonian:
    """Overlap, core Hamiltonian, a J/K builder and the electron counts of each spin."""

    def __init__(self, overlap, core_hamiltonian, two_body, n_alpha: int, n_beta: int,
                 nuclear_repulsion: float) -> None:
        n = two_body.n_ao
        device = two_body.device
        self.overlap = _tensor(overlap, device)
        self.core = _tensor(core_hamiltonian, device)
        if tuple(self.overlap.shape) != (n, n) or tuple(self.core.shape) != (n, n):
            raise InputError("the one-electron matrices act on the integrals' atomic orbitals")
        if not (0 <= n_beta and 0 <= n_alpha <= n and n_beta <= n):
            raise InputError("0 <= n_alpha, n_beta <= n_ao is required")
        self.two_body = two_body
        self.n_alpha, self.n_beta = int(n_alpha), int(n_beta)
        self.nuclear_repulsion = float(nuclear_repulsion)
        self.n_ao = n
        self.device = device
        self.dense = Dense(two_body._ctx, device)

    def fock(self, densities, orbitals=(None, None)):
        """``((F_a, F_b), terms)`` with ``terms`` = ``[tr(D h), tr(D J) / 2, -(tr(D_a K_a) +
        tr(D_b K_b)) / 2]`` on the device; ``orbitals[s]`` (``D_s = C_s C_s^T``) selects the
        occupied-form exchange of that spin."""
        torch = _torch()
        dn = self.dense
        total = dn.combine(1.0, densities[0], 1.0, densities[1])
        j = self.two_body.coulomb(total)
        ks = []
        for d, c in zip(densities, orbitals):
            ks.append(self.two_body.exchange_occupied(c) if c is not None else self.two_body.exchange(d))
        focks = tuple(dn.combine(1.0, self.core, 1.0, j, -1.0, k) for k in ks)
        parts = torch.empty(4, dtype=torch.float64, device=self.device)
        dn.dot(total, self.core, parts[0:1])
        dn.dot(total, j, parts[1:2])
        dn.dot(densities[0], ks[0], parts[2:3])
        dn.dot(densities[1], ks[1], parts[3:4])
        terms = torch.stack((parts[0], 0.5 * parts[1], -0.5 * (parts[2] + parts[3])))
        return focks, terms


@dataclass
class UhfResult:
    energies: EnergyTerms
    mo_coefficients: tuple
    mo_energies: tuple
    mo_occupations: tuple
    density: tuple
    fock: tuple
    spin_square: float
    converged: bool
    history: list = field(default_factory=list)
    seconds: float = 0.0

    @property
    def energy(self) -> float:
        return self.energies.total

    @property
    def n_iterations(self) -> int:
        return len(self.history)


def _aufbau(dn, c, count: int):
    c_occ = c[:, :count].contiguous()
    if count == 0:
        torch = _torch()
        return torch.zeros((c.shape[0], c.shape[0]), dtype=torch.float64, device=c.device), c_occ
    return dn.gemm(c_occ, c_occ, tb=True), c_occ


def spin_square(dn, d_alpha, d_beta, overlap, n_alpha: int, n_beta: int) -> float:
    """``<S^2>`` of a collinear determinant from its spin densities."""
    sz = 0.5 * (n_alpha - n_beta)
    a = dn.gemm(d_alpha, overlap)
    b = dn.gemm(d_beta, overlap)
    return sz * (sz + 1.0) + n_beta - float((a * b.T).sum().item())


def solve_uhf(hamiltonian: UnrestrictedHamiltonian, *, options: ScfOptions | None = None) -> UhfResult:
    """Converge the unrestricted state from the spin-averaged core guess.

    Raises ConvergenceError after ``max_iterations`` Fock builds (its ``last``
    holds the result) and BackendError when a diagonalisation fails.
    """
    torch = _torch()
    options = ScfOptions() if options is None else options
    if options.precision != "fp64":
        raise InputError("the unrestricted driver runs in fp64")
    if options.level_shift or options.damping:
        raise InputError("the unrestricted driver has no level shift or damping")
    t0 = time.perf_counter()
    setup = _Setup(hamiltonian, options)
    dn = hamiltonian.dense
    dn.reset_eig_failures()
    counts = (hamiltonian.n_alpha, hamiltonian.n_beta)
    if max(counts) > setup.kept:
        raise InputError(f"{max(counts)} occupied orbitals do not fit {setup.kept} retained functions")
    e, c = setup.solve(hamiltonian.core)
    occ = fill_degenerate_groups(e.cpu().numpy(), 0.5 * (counts[0] + counts[1]))
    scaled = dn.gemm(c, torch.diag(torch.from_numpy(occ)).to(c.device).contiguous())
    d0 = dn.gemm(scaled, c, tb=True)
    densities = (d0, d0.clone())
    orbitals = (None, None)
    focks_hist: list = []
    errors_hist: list = []
    overlaps = np.zeros((0, 0))
    history: list[ScfIteration] = []
    converged = False
    focks = None
    for index in range(options.max_iterations):
        t_it = time.perf_counter()
        focks, terms = hamiltonian.fock(densities, orbitals)
        grads = tuple(setup.gradient(f, d) for f, d in zip(focks, densities))
        focks_hist = (focks_hist + [focks])[-options.diis_size:]
        errors_hist = (errors_hist + [grads])[-options.diis_size:]
        m = len(errors_hist)
        scalars = torch.empty(4 + 2 * m, dtype=torch.float64, device=hamiltonian.device)
        scalars[0:3] = terms
        for i, (ga, gb) in enumerate(errors_hist):
            dn.dot(grads[0], ga, scalars[3 + 2 * i:4 + 2 * i])
            dn.dot(grads[1], gb, scalars[4 + 2 * i:5 + 2 * i])
        scalars[3 + 2 * m:4 + 2 * m] = dn.eig_failures
        host = scalars.cpu().numpy()
        if host[3 + 2 * m] != 0.0:
            raise BackendError("a Fock matrix could not be diagonalised (dsyevd reported a failure)")
        row = host[3:3 + 2 * m:2] + host[4:4 + 2 * m:2]
        kept_ov = overlaps[-(m - 1):, -(m - 1):] if m > 1 else np.zeros((0, 0))
        overlaps = np.zeros((m, m))
        overlaps[:m - 1, :m - 1] = kept_ov
        overlaps[m - 1, :] = row
        overlaps[:, m - 1] = row
        energies = EnergyTerms(hamiltonian.nuclear_repulsion, float(host[0]), float(host[1]), float(host[2]))
        gnorm = math.sqrt(max(row[m - 1], 0.0))
        total = energies.total
        change = total - history[-1].energy if history else math.inf
        history.append(ScfIteration(index, total, change, gnorm, "fp64", time.perf_counter() - t_it))
        if abs(change) < options.energy_tolerance and gnorm < options.gradient_tolerance:
            converged = True
            break
        if m == 1:
            extrapolated = focks_hist[0]
        else:
            coef = solve_diis(overlaps)
            extrapolated = []
            for s in range(2):
                acc = dn.combine(float(coef[0]), focks_hist[0][s])
                for ci, f in zip(coef[1:], focks_hist[1:]):
                    acc = dn.combine(1.0, acc, float(ci), f[s])
                extrapolated.append(acc)
        new_d, new_c = [], []
        for s in range(2):
            _, cs = setup.solve(extrapolated[s])
            d, c_occ = _aufbau(dn, cs, counts[s])
            new_d.append(d)
            new_c.append(c_occ)
        densities, orbitals = tuple(new_d), tuple(new_c)
    final = focks if converged else focks_hist[-1]
    mo_e, mo_c, occs, dens, orbs = [], [], [], [], []
    for s in range(2):
        es, cs = setup.solve(final[s])
        d, c_occ = _aufbau(dn, cs, counts[s])
        occ = np.zeros(setup.kept)
        occ[:counts[s]] = 1.0
        mo_e.append(es)
        mo_c.append(cs)
        occs.append(occ)
        dens.append(d)
        orbs.append(c_occ)
    fock, terms = hamiltonian.fock(tuple(dens), tuple(orbs))
    t = terms.cpu().numpy()
    energies = EnergyTerms(hamiltonian.nuclear_repulsion, float(t[0]), float(t[1]), float(t[2]))
    s2 = spin_square(dn, dens[0], dens[1], hamiltonian.overlap, counts[0], counts[1])
    result = UhfResult(energies, tuple(mo_c), tuple(mo_e), tuple(occs), tuple(dens), fock, s2, converged, history,
                       time.perf_counter() - t0)
    if not converged:
        last = history[-1]
        raise ConvergenceError(
            f"the SCF did not converge in {options.max_iterations} iterations (last energy change "
            f"{last.energy_change:.2e}, gradient {last.gradient_norm:.2e})",
            result,
        )
    return result
