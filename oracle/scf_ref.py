"""NumPy restatement of restricted Hartree-Fock with a J/K operator.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows, for the real restricted (RHF) structure:
* hafnium/scf/driver.py:154-213 (the loop), :239-246 (averaged core guess),
  :248-260 (orbital gradient ``X^T (F D S - S D F) X``), :307-329
  (``solve``/``next_density``), :385-395 (convergence, Frobenius norm with
  the spin weight 2 of a restricted block);
* hafnium/scf/hamiltonian.py:331-370 (``F = h + J[2 D_a] - K[D_a]``;
  ``E_1 = 2 tr(D_a h)``, ``E_J = tr(D_a J)``, ``E_K = -tr(D_a K)``, the
  structured trace of a restricted block weighting 2, structured/dense.py:
  241-280), coulomb_exchange.py:407-462 (J of the spin-traced density
  ``2 D_a``, K of the stored block ``D_a``);
* hafnium/scf/diis.py:27-106 (Pulay DIIS with the eigen-solve of the
  bordered system and the null space dropped at 1e-14);
* hafnium/scf/orthogonalization.py:16-42 (canonical orthogonaliser);
* hafnium/scf/occupation.py:289-370 (aufbau and degenerate-group filling);
* hafnium/integrals/density_fitting.py:97-107, :181-197 (fitted tensor
  ``M^-1/2 (P|uv)``, eigenvalues below 1e-10 of the largest dropped) and the
  Cholesky variant ``L^-1 (P|uv)`` with ``M = L L^T``, which gives the same
  ``B^T B`` and hence the same J and K.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from oracle import jk_ref

DEFAULT_METRIC_THRESHOLD = 1e-10
DEFAULT_OVERLAP_THRESHOLD = 1e-10
NULL_SPACE_TOLERANCE = 1e-14
FILL_TOLERANCE = 1e-12


def inverse_square_root(metric: np.ndarray, threshold: float = DEFAULT_METRIC_THRESHOLD) -> np.ndarray:
    """``M^-1/2`` dropping eigenvalues below ``threshold * max`` (density_fitting.py:181-197)."""
    w, v = np.linalg.eigh(metric)
    keep = w > threshold * np.max(w)
    inv = np.where(keep, 1.0 / np.sqrt(np.where(keep, w, 1.0)), 0.0)
    return (v * inv) @ v.T


def df_tensor(metric: np.ndarray, three_center: np.ndarray, method: str = "eigh") -> np.ndarray:
    """``B[aux, ao, ao]``: ``M^-1/2 (P|uv)`` (eigh) or ``L^-1 (P|uv)`` (cholesky)."""
    naux = metric.shape[0]
    flat = three_center.reshape(naux, -1)
    if method == "eigh":
        out = inverse_square_root(metric) @ flat
    elif method == "cholesky":
        lower = np.linalg.cholesky(metric)
        out = np.linalg.solve(lower, flat)  # triangular solve, dense routine is fine here
    else:
        raise ValueError(method)
    return out.reshape(three_center.shape)


def canonical_orthogonalizer(overlap: np.ndarray, threshold: float = DEFAULT_OVERLAP_THRESHOLD) -> np.ndarray:
    """``X = V s^-1/2`` over eigenpairs with ``s > threshold`` (orthogonalization.py:16-42)."""
    w, v = np.linalg.eigh(overlap)
    kept = int(np.sum(w > threshold))
    if kept == 0:
        raise ValueError("the overlap is not positive definite")
    n = overlap.shape[0]
    return v[:, n - kept:] / np.sqrt(w[n - kept:])[None, :]


def solve_diis(overlaps: np.ndarray) -> np.ndarray:
    """DIIS coefficients of the ``[m, m]`` error overlaps (diis.py:82-106)."""
    m = overlaps.shape[0]
    scale = np.max(np.diag(overlaps))
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


@dataclass
class Diis:
    """The last ``size`` Fock matrices and errors (diis.py:27-67)."""

    size: int = 8
    focks: list = field(default_factory=list)
    errors: list = field(default_factory=list)

    def push(self, fock: np.ndarray, error: np.ndarray) -> None:
        self.focks = (self.focks + [fock])[-self.size:]
        self.errors = (self.errors + [error])[-self.size:]

    def extrapolate(self) -> np.ndarray:
        if len(self.focks) == 1:
            return self.focks[0]
        # Real overlap of the restricted errors under the structured trace
        # (weight 2); the scale cancels in solve_diis.
        ov = np.array([[2.0 * float(np.sum(a * b)) for b in self.errors] for a in self.errors])
        c = solve_diis(ov)
        out = c[0] * self.focks[0]
        for ci, f in zip(c[1:], self.focks[1:]):
            out = out + ci * f
        return out


def fill_whole_levels(energies: np.ndarray, count: int) -> np.ndarray:
    """Aufbau: ones for the ``count`` lowest levels (occupation.py:321-339)."""
    occ = np.zeros(energies.shape[0])
    occ[np.argsort(energies, kind="stable")[:count]] = 1.0
    return occ


def fill_degenerate_groups(energies: np.ndarray, count: float, tolerance: float = 1e-6) -> np.ndarray:
    """Averaged filling of degenerate groups (occupation.py:342-370)."""
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
        capacity = float(len(group))
        share = min(remaining, capacity) / capacity
        occ[group] = share
        remaining = max(remaining - share * capacity, 0.0)
        start = end
    if not math.isclose(remaining, 0.0, abs_tol=FILL_TOLERANCE):
        raise ValueError("electrons left over")
    return occ


@dataclass
class RhfResult:
    energy: float
    one_electron: float
    coulomb: float
    exchange: float
    nuclear_repulsion: float
    mo_energies: np.ndarray
    mo_coefficients: np.ndarray
    occupations: np.ndarray
    density: np.ndarray  # D_alpha
    fock: np.ndarray
    converged: bool
    history: list


def fock_and_energies(h, j_of, k_of, density, e_nuc):
    """``F = h + J[2 D_a] - K[D_a]`` and the energy terms (hamiltonian.py:331-370)."""
    j = j_of(2.0 * density)
    k = k_of(density)
    fock = h + j - k
    e1 = 2.0 * float(np.sum(density * h.T))
    ej = 2.0 * float(np.sum(density * j.T)) / 2.0
    ek = -2.0 * float(np.sum(density * k.T)) / 2.0
    return fock, (e_nuc, e1, ej, ek)


def rhf(
    overlap: np.ndarray,
    core: np.ndarray,
    j_of,
    k_of,
    n_alpha: int,
    nuclear_repulsion: float,
    *,
    guess: np.ndarray | None = None,
    max_iterations: int = 100,
    energy_tolerance: float = 1e-10,
    gradient_tolerance: float = 1e-6,
    diis_size: int = 8,
    occupation: str = "aufbau",
) -> RhfResult:
    """Converge a closed-shell RHF state (driver.py:154-213, :307-367).

    ``occupation`` "aufbau" fills whole levels; "averaged" spreads ``n_alpha`` (possibly
    fractional) evenly over degenerate groups in every iteration, the reference's
    ``AveragedOccupation`` (occupation.py:342-370) used for the spherical atoms of its density
    guess (atoms.py:55-73)."""
    if occupation not in ("aufbau", "averaged"):
        raise ValueError("occupation is 'aufbau' or 'averaged'")
    fill = fill_whole_levels if occupation == "aufbau" else fill_degenerate_groups
    x = canonical_orthogonalizer(overlap)

    def solve(f):
        e, c = np.linalg.eigh(x.T @ f @ x)
        return e, x @ c

    if guess is None:
        e, c = solve(core)
        occ = fill_degenerate_groups(e, n_alpha)
        density = (c * occ[None, :]) @ c.T
    else:
        density = guess
    diis = Diis(diis_size)
    history = []
    converged = False
    fock = None
    for index in range(max_iterations):
        fock, terms = fock_and_energies(core, j_of, k_of, density, nuclear_repulsion)
        total = sum(terms)
        grad = x.T @ (fock @ density @ overlap - overlap @ density @ fock) @ x
        gnorm = math.sqrt(2.0 * float(np.sum(grad * grad)))
        change = total - history[-1][1] if history else math.inf
        history.append((index, total, change, gnorm))
        if abs(change) < energy_tolerance and gnorm < gradient_tolerance:
            converged = True
            break
        diis.push(fock, grad)
        e, c = solve(diis.extrapolate())
        occ = fill(e, n_alpha)
        density = (c * occ[None, :]) @ c.T
    final = fock if converged else diis.focks[-1]
    e, c = solve(final)
    occ = fill(e, n_alpha)
    density = (c * occ[None, :]) @ c.T
    fock, terms = fock_and_energies(core, j_of, k_of, density, nuclear_repulsion)
    return RhfResult(
        energy=sum(terms),
        one_electron=terms[1],
        coulomb=terms[2],
        exchange=terms[3],
        nuclear_repulsion=terms[0],
        mo_energies=e,
        mo_coefficients=c,
        occupations=occ,
        density=density,
        fock=fock,
        converged=converged,
        history=history,
    )


def fitted_operators(b: np.ndarray):
    """``(J(D), K(D))`` callables of the fitted tensor (density form)."""
    return (lambda d: jk_ref.coulomb(b, d)), (lambda d: jk_ref.exchange(b, d))


def exact_operators(eri: np.ndarray):
    """``(J(D), K(D))`` of full four-centre integrals (coulomb_exchange.py:113-121)."""
    return (
        lambda d: np.einsum("uvwx,wx->uv", eri, d),
        lambda d: np.einsum("uwxv,wx->uv", eri, d),
    )


def nuclear_repulsion(charges, positions_bohr) -> float:
    """``sum_{A<B} Z_A Z_B / |R_A - R_B|`` (molecule/molecule.py:45-56)."""
    total = 0.0
    pos = np.asarray(positions_bohr, dtype=float)
    for a in range(len(charges)):
        for b in range(a):
            total += charges[a] * charges[b] / float(np.linalg.norm(pos[a] - pos[b]))
    return total


@dataclass
class UhfResult:
    energy: float
    mo_energies: tuple  # (alpha, beta)
    mo_coefficients: tuple
    occupations: tuple
    density: tuple  # (D_alpha, D_beta)
    spin_square: float
    converged: bool
    history: list


def spin_square(d_alpha: np.ndarray, d_beta: np.ndarray, overlap: np.ndarray, n_alpha: float, n_beta: float) -> float:
    """``<S^2> = S_z (S_z + 1) + n_beta - tr(D_a S D_b S)`` of a collinear determinant."""
    sz = 0.5 * (n_alpha - n_beta)
    return sz * (sz + 1.0) + n_beta - float(np.sum((d_alpha @ overlap) * (d_beta @ overlap).T))


def uhf(
    overlap: np.ndarray,
    core: np.ndarray,
    j_of,
    k_of,
    n_alpha: int,
    n_beta: int,
    nuclear_repulsion: float,
    *,
    max_iterations: int = 100,
    energy_tolerance: float = 1e-10,
    gradient_tolerance: float = 1e-6,
    diis_size: int = 8,
) -> UhfResult:
    """Converge an unrestricted (collinear) determinant (driver.py:154-213 for the collinear
    structure).

    * Start: the spin-averaged core guess, ``n_electrons / 2`` per spin with the averaged
      filling (driver.py:216-246).
    * Fock: ``F_s = h + J[D_a + D_b] - K[D_s]`` (coulomb_exchange.py:407-462: J of the
      spin trace on both blocks, K per stored block).
    * Energy: ``tr((D_a + D_b) h) + tr((D_a + D_b) J) / 2 - (tr(D_a K_a) + tr(D_b K_b)) / 2``
      (hamiltonian.py:331-370 with weight 1 per collinear block).
    * Gradient ``X^T (F_s D_s S - S D_s F_s) X`` per block; its norm and the DIIS overlaps sum
      over both blocks (driver.py:248-260, :385-395; diis.py:55-79).
    """
    x = canonical_orthogonalizer(overlap)

    def solve(f):
        e, c = np.linalg.eigh(x.T @ f @ x)
        return e, x @ c

    e, c = solve(core)
    half = 0.5 * (n_alpha + n_beta)
    occ = fill_degenerate_groups(e, half)
    d0 = (c * occ[None, :]) @ c.T
    dens = [d0, d0.copy()]
    counts = (n_alpha, n_beta)

    def build(dens):
        total = dens[0] + dens[1]
        j = j_of(total)
        ks = [k_of(d) for d in dens]
        focks = [core + j - k for k in ks]
        e1 = float(np.sum(total * core.T))
        ej = 0.5 * float(np.sum(total * j.T))
        ek = -0.5 * sum(float(np.sum(d * k.T)) for d, k in zip(dens, ks))
        return focks, nuclear_repulsion + e1 + ej + ek

    focks_hist, errors_hist, history = [], [], []
    converged = False
    for index in range(max_iterations):
        focks, total = build(dens)
        grads = [x.T @ (f @ d @ overlap - overlap @ d @ f) @ x for f, d in zip(focks, dens)]
        gnorm = math.sqrt(sum(float(np.sum(g * g)) for g in grads))
        change = total - history[-1][1] if history else math.inf
        history.append((index, total, change, gnorm))
        if abs(change) < energy_tolerance and gnorm < gradient_tolerance:
            converged = True
            break
        focks_hist = (focks_hist + [focks])[-diis_size:]
        errors_hist = (errors_hist + [grads])[-diis_size:]
        m = len(focks_hist)
        if m == 1:
            extrapolated = focks_hist[0]
        else:
            ov = np.array([[sum(float(np.sum(ga * gb)) for ga, gb in zip(a, b)) for b in errors_hist]
                           for a in errors_hist])
            coef = solve_diis(ov)
            extrapolated = [sum(ci * f[s] for ci, f in zip(coef, focks_hist)) for s in range(2)]
        dens = []
        for s in range(2):
            e, c = solve(extrapolated[s])
            occ = fill_whole_levels(e, counts[s])
            dens.append((c * occ[None, :]) @ c.T)
    final = focks if converged else focks_hist[-1]
    energies, coefficients, occupations, dens = [], [], [], []
    for s in range(2):
        e, c = solve(final[s])
        occ = fill_whole_levels(e, counts[s])
        energies.append(e)
        coefficients.append(c)
        occupations.append(occ)
        dens.append((c * occ[None, :]) @ c.T)
    _, total = build(dens)
    return UhfResult(
        energy=total,
        mo_energies=tuple(energies),
        mo_coefficients=tuple(coefficients),
        occupations=tuple(occupations),
        density=tuple(dens),
        spin_square=spin_square(dens[0], dens[1], overlap, n_alpha, n_beta),
        converged=converged,
        history=history,
    )
