"""Restricted Hartree-Fock on the GPU around the MI355X DF-J/K build.

Mirrors, for the real restricted (RHF) structure, the reference's
``solve_scf`` loop (hafnium/scf/driver.py:154-213), its orbital gradient and
diagonalisation (:248-260, :307-329), its final consistent rebuild
(:331-367) and convergence test (:385-395), ``Hamiltonian.fock`` and the
energy terms (hafnium/scf/hamiltonian.py:40-66, :331-370), ``Diis``
(hafnium/scf/diis.py:27-106), the canonical orthogonaliser
(hafnium/scf/orthogonalization.py:16-42) and the aufbau / averaged
occupations (hafnium/scf/occupation.py:289-370). ``ConvergenceError`` carries
the last result as ``last`` (hafnium/errors.py:33-43).

Added for MI355X: a precision policy. J and K run in ``options.precision``
("bf16x3", "fp32" or "fp64") while the orbital-gradient norm (the DIIS
error) is at or above ``options.switch_at(precision)`` (the switch
threshold, raised to the precision's floor) and still falling
(``mixed_phase_stalled``), then in FP64 for every remaining iteration;
convergence is only declared on FP64 iterations, and
the final rebuild is FP64, so the converged energy is that of the FP64
operator. The Fock diagonalisation, DIIS, the gradient and all energy sums
stay FP64. The K of the loop uses the occupied orbitals of the previous
diagonalisation (``exchange_occupied``); only a guess density without
orbitals goes through the density-form exchange.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from hafnium_amd.errors import BackendError, ConvergenceError, InputError
from hafnium_amd.linalg import Dense

DEFAULT_OVERLAP_THRESHOLD = 1e-10
NULL_SPACE_TOLERANCE = 1e-14
FILL_TOLERANCE = 1e-12


def _torch():
    import torch

    return torch


@dataclass(frozen=True)
class ScfOptions:
    """Tolerances and limits (driver.py:38-84) and the precision policy."""

    max_iterations: int = 100
    energy_tolerance: float = 1e-10
    gradient_tolerance: float = 1e-6
    diis_size: int = 8
    level_shift: float = 0.0
    damping: float = 0.0
    overlap_threshold: float = DEFAULT_OVERLAP_THRESHOLD
    precision: str = "fp64"
    switch_threshold: float = 1e-5
    # Gradient norms below which a reduced-precision build no longer tells the FP64 gradient:
    # about three times the distance between the reduced-precision and the FP64 fixed points
    # measured on water clusters of up to 12 waters (FP32 3e-5, BF16x3 3.6e-4).
    switch_floor_fp32: float = 1e-4
    switch_floor_bf16x3: float = 1e-3
    reset_diis_on_switch: bool = True
    stall_window: int = 2
    stall_ratio: float = 0.7

    def __post_init__(self) -> None:
        if self.max_iterations < 1:
            raise InputError(f"max_iterations = {self.max_iterations}: the loop needs room for one iteration")
        if self.energy_tolerance <= 0.0 or self.gradient_tolerance <= 0.0:
            raise InputError("the energy and gradient tolerances have to be positive numbers")
        if self.diis_size < 1:
            raise InputError(f"diis_size = {self.diis_size}: DIIS extrapolates over one stored Fock matrix or more")
        if self.level_shift < 0.0:
            raise InputError(f"a negative level shift ({self.level_shift}) would lower the virtual levels")
        if not 0.0 <= self.damping < 1.0:
            raise InputError(f"damping = {self.damping}: the mixing weight of the previous density lies in [0, 1)")
        if self.overlap_threshold <= 0.0:
            raise InputError(f"the overlap threshold {self.overlap_threshold} must exceed zero")
        if self.precision not in ("fp64", "fp32", "bf16x3"):
            raise InputError("precision is fp64, fp32 or bf16x3")
        if self.stall_window < 1 or not 0.0 < self.stall_ratio <= 1.0:
            raise InputError("stall_window >= 1 and 0 < stall_ratio <= 1 are required")
        if self.switch_threshold <= 0.0 or self.switch_floor_fp32 < 0.0 or self.switch_floor_bf16x3 < 0.0:
            raise InputError("the switch threshold is positive and the floors non-negative")

    def switch_at(self, precision: str) -> float:
        """Gradient norm below which a ``precision`` iteration hands over to FP64: the switch
        threshold, or that precision's floor where the floor is higher."""
        floor = {"fp32": self.switch_floor_fp32, "bf16x3": self.switch_floor_bf16x3}.get(precision, 0.0)
        return max(self.switch_threshold, floor)


def mixed_phase_stalled(gradient_norms, window: int, ratio: float) -> bool:
    """True when the last ``window`` reduced-precision iterations did not bring the
    orbital gradient below ``ratio`` times the best norm before them.

    The J/K error of a reduced arithmetic sets a floor under the gradient
    (about 1e-4 for BF16x3 on small water clusters, above the 1e-5 switch
    threshold); once DIIS stops making progress against that floor, further
    reduced-precision iterations are wasted.
    """
    g = list(gradient_norms)
    if len(g) < window + 2:
        return False
    return min(g[-window:]) > ratio * min(g[:-window])


@dataclass(frozen=True)
class EnergyTerms:
    """``E = E_nuc + tr(D h) + tr(D J)/2 - tr(D K)/2`` (hamiltonian.py:40-66)."""

    nuclear_repulsion: float
    one_electron: float
    coulomb: float
    exchange: float

    @property
    def electronic(self) -> float:
        return self.one_electron + self.coulomb + self.exchange

    @property
    def total(self) -> float:
        return self.nuclear_repulsion + self.electronic


@dataclass(frozen=True)
class ScfIteration:
    index: int
    energy: float
    energy_change: float
    gradient_norm: float
    precision: str
    seconds: float


@dataclass
class ScfResult:
    energies: EnergyTerms
    mo_coefficients: object
    mo_energies: object
    mo_occupations: object
    density: object
    fock: object
    converged: bool
    history: list = field(default_factory=list)
    seconds: float = 0.0

    @property
    def energy(self) -> float:
        return self.energies.total

    @property
    def n_iterations(self) -> int:
        return len(self.history)


def solve_diis(overlaps: np.ndarray) -> np.ndarray:
    """DIIS coefficients (diis.py:82-106): eigen-solve of the bordered system, null space dropped.

    A host computation on an at most (size+1)^2 matrix, like the reference's.
    """
    m = overlaps.shape[0]
    scale = float(np.max(np.diag(overlaps)))
    overlaps = overlaps / (scale if scale > 0.0 else 1.0)
    system = np.zeros((m + 1, m + 1))
    system[:m, :m] = overlaps
    system[:m, m] = -1.0
    system[m, :m] = -1.0
    right = np.zeros(m + 1)
    right[m] = -1.0
    w, v = np.linalg.eigh(system)
    keep = np.abs(w) > NULL_SPACE_TOLERANCE * np.max(np.abs(w))
    inv = np.where(keep, 1.0 / np.where(keep, w, 1.0), 0.0)
    return (v @ (inv * (v.T @ right)))[:m]


def fill_degenerate_groups(energies: np.ndarray, count: float, tolerance: float = 1e-6) -> np.ndarray:
    """``AveragedOccupation`` filling (occupation.py:342-370), for the core guess."""
    order = np.argsort(energies, kind="stable")
    vals = energies[order]
    spread = max(1.0, float(vals[-1] - vals[0])) if len(vals) else 1.0
    occ = np.zeros(len(energies))
    remaining = float(count)
    start = 0
    while start < len(order):
        end = start + 1
        while end < len(order) and vals[end] - vals[end - 1] <= tolerance * spread:
            end += 1
        group = order[start:end]
        share = min(remaining, float(len(group))) / float(len(group))
        occ[group] = share
        remaining = max(remaining - share * len(group), 0.0)
        start = end
    if not math.isclose(remaining, 0.0, abs_tol=FILL_TOLERANCE):
        raise InputError(f"cannot place {count} electrons in {len(energies)} orbitals")
    return occ


class RestrictedHamiltonian:
    """Overlap, core Hamiltonian, a J/K builder and the electron count of a closed shell.

    The RHF case of the reference's ``Hamiltonian`` (hamiltonian.py:69-130):
    ``two_body`` is any object with ``coulomb``/``exchange`` and, for the fast
    path, ``exchange_occupied`` (``MI355XCoulombExchange``).
    """

    def __init__(self, overlap, core_hamiltonian, two_body, n_alpha: int, nuclear_repulsion: float) -> None:
        torch = _torch()
        n = two_body.n_ao
        device = two_body.device
        self.overlap = torch.as_tensor(np.asarray(overlap) if isinstance(overlap, np.ndarray) else overlap,
                                       dtype=torch.float64).to(device).contiguous()
        self.core = torch.as_tensor(np.asarray(core_hamiltonian) if isinstance(core_hamiltonian, np.ndarray)
                                    else core_hamiltonian, dtype=torch.float64).to(device).contiguous()
        if tuple(self.overlap.shape) != (n, n) or tuple(self.core.shape) != (n, n):
            raise InputError("the one-electron matrices act on the integrals' atomic orbitals")
        if n_alpha < 0 or n_alpha > n:
            raise InputError("0 <= n_alpha <= n_ao is required")
        self.two_body = two_body
        # Fractional counts are allowed for the averaged occupation (atoms).
        self.n_alpha = int(n_alpha) if float(n_alpha).is_integer() else float(n_alpha)
        self.nuclear_repulsion = float(nuclear_repulsion)
        self.n_ao = n
        self.device = device
        self.dense = Dense(two_body._ctx, device)

    def fock(self, density, orbitals=None, precision: str = "fp64"):
        """``F = h + J[2 D] - K[D]`` and the four energy terms as a device vector.

        ``orbitals`` (``[n, n_alpha]`` with ``D = C C^T``) selects the
        occupied-form exchange. Returns ``(F, terms)`` with ``terms`` =
        ``[2 tr(D h), tr(D J[2D]), -tr(D K)]`` on the device.
        """
        torch = _torch()
        dn = self.dense
        if orbitals is not None:
            # D = C C^T: J and K in one call (coulomb_exchange.py:397-404).
            j1, k = self.two_body.jk_occupied(orbitals, precision=precision)
        else:
            j1 = self.two_body.coulomb(density, precision=precision)
            k = self.two_body.exchange(density, precision=precision)
        fock = dn.combine(1.0, self.core, 2.0, j1, -1.0, k)
        terms = torch.empty(3, dtype=torch.float64, device=self.device)
        dn.dot(density, self.core, terms[0:1])
        dn.dot(density, j1, terms[1:2])
        dn.dot(density, k, terms[2:3])
        return fock, terms


def _energy_terms(hamiltonian, host_terms) -> EnergyTerms:
    d_h, d_j, d_k = (float(v) for v in host_terms)
    return EnergyTerms(hamiltonian.nuclear_repulsion, 2.0 * d_h, 2.0 * d_j, -d_k)


class _Setup:
    """Orthogonaliser and the diagonalisation step (driver.py:263-329)."""

    def __init__(self, hamiltonian: RestrictedHamiltonian, options: ScfOptions) -> None:
        self.h = hamiltonian
        self.options = options
        dn = hamiltonian.dense
        w, v = dn.eigh(hamiltonian.overlap)
        w_host = w.cpu().numpy()
        if dn.info() != 0:
            raise BackendError(f"the overlap could not be diagonalised (dsyevd info {dn.info()})")
        kept = int(np.sum(w_host > options.overlap_threshold))
        if kept == 0:
            raise InputError("the overlap is not positive definite")
        n = hamiltonian.n_ao
        torch = _torch()
        inv_root = torch.diag(torch.from_numpy(1.0 / np.sqrt(w_host[n - kept:]))).to(hamiltonian.device)
        # X = V s^-1/2 over the retained eigenpairs (orthogonalization.py:16-42).
        self.x = dn.gemm(v[:, n - kept:].contiguous(), inv_root.contiguous())
        self.kept = kept

    def solve(self, fock):
        """Canonical orbitals of ``fock``: ``(energies, C [n, kept])``."""
        dn = self.h.dense
        xf = dn.gemm(self.x, fock, ta=True)
        reduced = dn.gemm(xf, self.x)
        e, c_red = dn.eigh(reduced)
        return e, dn.gemm(self.x, c_red)

    def gradient(self, fock, density):
        """``X^T (F D S - S D F) X`` (driver.py:248-260)."""
        dn = self.h.dense
        fd = dn.gemm(fock, density)
        fds = dn.gemm(fd, self.h.overlap)
        comm = dn.antisymmetrize(fds)
        xc = dn.gemm(self.x, comm, ta=True)
        return dn.gemm(xc, self.x)


def _density(dense, c, occupations=None, n_alpha=None):
    """``D = C diag(occ) C^T`` (structured/density.py:16-41); aufbau when ``n_alpha`` is given."""
    torch = _torch()
    if n_alpha is not None:
        c_occ = c[:, :n_alpha].contiguous()
        return dense.gemm(c_occ, c_occ, tb=True), c_occ
    occ = torch.diag(torch.from_numpy(np.asarray(occupations, dtype=np.float64))).to(c.device).contiguous()
    scaled = dense.gemm(c, occ)
    return dense.gemm(scaled, c, tb=True), None


def solve_scf(hamiltonian: RestrictedHamiltonian, *, guess=None, options: ScfOptions | None = None,
              occupation: str = "aufbau", guess_orbitals=None, first_build=None) -> ScfResult:
    """Converge the closed-shell state (driver.py:154-213) with the precision policy.

    ``guess``: a density ``[n, n]`` (D_alpha) or ``None`` for the averaged
    core-Hamiltonian guess (driver.py:216-246). ``occupation``: "aufbau"
    (``AufbauOccupation``) or "averaged" (``AveragedOccupation``: degenerate
    frontier groups share the electrons, for spherical atoms;
    occupation.py:130-159).
    ``guess_orbitals`` (``[n, n_alpha]`` with guess = C C^T) selects the
    occupied-form exchange for the first build; ``first_build`` =
    ``(fock, terms)`` of ``hamiltonian.fock(guess, ...)`` already computed by
    the caller is used as the first iteration's build.
    Raises ConvergenceError after ``max_iterations`` Fock builds; its
    ``last`` holds the result.
    """
    torch = _torch()
    options = ScfOptions() if options is None else options
    if occupation not in ("aufbau", "averaged"):
        raise InputError("occupation is 'aufbau' or 'averaged'")
    t0 = time.perf_counter()
    setup = _Setup(hamiltonian, options)
    dn = hamiltonian.dense
    dn.reset_eig_failures()
    n_alpha = hamiltonian.n_alpha
    if n_alpha > setup.kept:
        raise InputError(f"{n_alpha} occupied orbitals do not fit {setup.kept} retained functions")
    if occupation == "aufbau" and not isinstance(n_alpha, int):
        raise InputError(f"aufbau filling needs an integer count, got {n_alpha}")

    def occupy(e, c):
        if occupation == "averaged":
            occ = fill_degenerate_groups(e.cpu().numpy(), n_alpha)
            return _density(dn, c, occupations=occ)
        return _density(dn, c, n_alpha=n_alpha)

    orbitals = None
    if guess is None:
        e, c = setup.solve(hamiltonian.core)
        occ = fill_degenerate_groups(e.cpu().numpy(), n_alpha)
        density, _ = _density(dn, c, occupations=occ)
    else:
        density = torch.as_tensor(np.asarray(guess) if isinstance(guess, np.ndarray) else guess,
                                  dtype=torch.float64).to(hamiltonian.device).contiguous()
        if tuple(density.shape) != (hamiltonian.n_ao, hamiltonian.n_ao):
            from hafnium_amd.errors import StructureError

            raise StructureError(f"the guess acts on {tuple(density.shape)}, the Hamiltonian on {hamiltonian.n_ao}")
        if guess_orbitals is not None:
            orbitals = torch.as_tensor(guess_orbitals, dtype=torch.float64).to(hamiltonian.device).contiguous()
            if tuple(orbitals.shape) != (hamiltonian.n_ao, n_alpha):
                raise InputError("guess orbitals are [n_ao, n_alpha]")
    if first_build is not None and guess is None:
        raise InputError("a first build needs the guess density it was built from")
    precision = options.precision
    focks: list = []
    errors: list = []
    # 2 tr(e_i e_j) of the stored errors, kept across iterations: each iteration's transfer
    # brings only the newest row.
    overlaps = np.zeros((0, 0))
    history: list[ScfIteration] = []
    converged = False
    fock = None
    previous = None
    for index in range(options.max_iterations):
        t_it = time.perf_counter()
        if index == 0 and first_build is not None:
            fock, terms = first_build
        else:
            fock, terms = hamiltonian.fock(density, orbitals, precision)
        grad = setup.gradient(fock, density)
        # One host transfer per iteration: energy terms, the newest error's overlaps with the
        # stored ones (its own norm last), and the count of failed diagonalisations.
        focks = (focks + [fock])[-options.diis_size:]
        errors = (errors + [grad])[-options.diis_size:]
        m = len(errors)
        scalars = torch.empty(5 + m, dtype=torch.float64, device=hamiltonian.device)
        scalars[0:3] = terms
        for i, e_i in enumerate(errors):
            dn.dot(grad, e_i, scalars[3 + i:4 + i])
        scalars[3 + m:4 + m] = dn.eig_failures
        host = scalars.cpu().numpy()
        if host[3 + m] != 0.0:
            raise BackendError("the Fock matrix could not be diagonalised (dsyevd reported a failure)")
        row = 2.0 * host[3:3 + m]
        kept_ov = overlaps[-(m - 1):, -(m - 1):] if m > 1 else np.zeros((0, 0))
        overlaps = np.zeros((m, m))
        overlaps[:m - 1, :m - 1] = kept_ov
        overlaps[m - 1, :] = row
        overlaps[:, m - 1] = row
        energies = _energy_terms(hamiltonian, host[0:3])
        gnorm = math.sqrt(max(row[m - 1], 0.0))
        total = energies.total
        change = total - history[-1].energy if history else math.inf
        history.append(ScfIteration(index, total, change, gnorm, precision, time.perf_counter() - t_it))
        if precision == "fp64" and abs(change) < options.energy_tolerance and gnorm < options.gradient_tolerance:
            converged = True
            break
        mixed_norms = [row.gradient_norm for row in history if row.precision != "fp64"]
        if precision != "fp64" and (gnorm < options.switch_at(precision)
                                    or mixed_phase_stalled(mixed_norms, options.stall_window, options.stall_ratio)
                                    or len(mixed_norms) >= options.max_iterations // 2):
            precision = "fp64"
            if options.reset_diis_on_switch:
                focks, errors = [fock], [grad]
                overlaps = overlaps[-1:, -1:]
        # DIIS extrapolation (diis.py:55-79) from the overlaps of the stored errors.
        if len(focks) == 1:
            extrapolated = focks[0]
        else:
            coeff = solve_diis(overlaps)
            extrapolated = dn.combine(float(coeff[0]), focks[0])
            for ci, f in zip(coeff[1:], focks[1:]):
                extrapolated = dn.combine(1.0, extrapolated, float(ci), f)
        if options.level_shift:
            sds = dn.gemm(dn.gemm(hamiltonian.overlap, density), hamiltonian.overlap)
            virtual = dn.combine(1.0, hamiltonian.overlap, -1.0, sds)
            extrapolated = dn.combine(1.0, extrapolated, options.level_shift, virtual)
        previous = density
        e, c = setup.solve(extrapolated)
        density, orbitals = occupy(e, c)
        if options.damping:
            density = dn.combine(1.0 - options.damping, density, options.damping, previous)
            orbitals = None
    final_fock = fock if converged else focks[-1]
    e, c = setup.solve(final_fock)
    density, orbitals = occupy(e, c)
    fock, terms = hamiltonian.fock(density, orbitals, "fp64")
    energies = _energy_terms(hamiltonian, terms.cpu().numpy())
    if occupation == "averaged":
        occupations = fill_degenerate_groups(e.cpu().numpy(), n_alpha)
    else:
        occupations = np.zeros(setup.kept)
        occupations[:n_alpha] = 1.0
    result = ScfResult(energies, c, e, occupations, density, fock, converged, history, time.perf_counter() - t0)
    if not converged:
        last = history[-1]
        raise ConvergenceError(
            f"the SCF did not converge in {options.max_iterations} iterations (last energy change "
            f"{last.energy_change:.2e}, gradient {last.gradient_norm:.2e})",
            result,
        )
    return result
