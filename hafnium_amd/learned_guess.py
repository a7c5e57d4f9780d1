"""Learned SCF starting guess: fitted-density coefficients predicted per atom.

The model predicts, for every atom, the coefficients of the Coulomb-fitted
electron density ``rho(r) = sum_P d_P chi_P(r)`` on that atom's auxiliary
functions, ``d = M^-1 (P|uv) (2 D_alpha)`` (M the Coulomb metric), from
descriptors of the atom's environment. Waters are rigid, so each atom uses
the frame of its own water (origin O, z along the H-O-H bisector, x along
H1 - H2): descriptors are computed in that frame, and the predicted
coefficients of l > 0 functions are rotated back to the laboratory frame
with the real-solid-harmonic rotation matrices of ``rotation_matrices``.

Descriptors: for each neighbour element (H, O) within 6 Å, 16 radial
Gaussians with centres from 0.5 to 6 Å, weighted by a cosine cutoff and by
the real solid harmonics (l <= 2) of the neighbour direction in the local
frame: 2 x 16 x 9 = 288 numbers per atom.

Network: one MLP per element, 288 -> 256 -> 256 -> n_coef (SiLU), BF16
weights with FP32 accumulation, run and trained by the library's own
kernels (``hafnium_amd.mlp``). Targets are whitened with the element's
one-centre Coulomb metric G = U^T U, y = U (c_local - mean) / scale, so the
squared error is the Coulomb self-energy of each atom's density error.

Guess: ``F0 = h + J[d] - K[D_sad]`` (Coulomb from the predicted density,
exchange from the superposition of atomic densities); the aufbau density
of F0 starts the SCF. When its energy is above the energy of the atomic
guess, the SCF starts from the atomic guess instead (the fallback), exactly
as it would without the network.

Training data: converged FP64 SCF densities of seeded water clusters
produced by this library's own GPU SCF (``generate_dataset``).

No counterpart in the reference (its guesses are the core Hamiltonian and
the superposition of atomic densities, hafnium/scf/driver.py:216-246 and
hafnium/scf/atoms.py:76-119).
"""

from __future__ import annotations

import argparse
import functools
import math
import os
import sys
import time
from dataclasses import dataclass

import numpy as np

from hafnium_amd.errors import InputError
from hafnium_amd.molecule import ANGSTROM_IN_BOHR, BasisOnAtoms, Molecule, load_basis, water_cluster

N_RADIAL = 16
R_MIN, R_MAX = 0.5, 6.0  # Å
L_DESCRIPTOR = 2
ELEMENTS = (1, 8)
N_FEATURES = len(ELEMENTS) * N_RADIAL * (L_DESCRIPTOR + 1) ** 2
HIDDEN = (256, 256)
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "learned_guess_water.npz")


def _torch():
    import torch

    return torch


# ---------------------------------------------------------------- solid harmonics

def _cart_list(l):
    out = []
    for i in range(l, -1, -1):
        for j in range(l - i, -1, -1):
            out.append((i, j, l - i - j))
    return out


def _sph_order(l):
    return [1, -1, 0] if l == 1 else list(range(-l, l + 1))


def _binom(n, k):
    return math.comb(n, k) if 0 <= k <= n else 0


def cart_to_sph(l: int) -> np.ndarray:
    """Real solid harmonics as Cartesian monomial sums, in the integral engine's order
    (p: x, y, z; l != 1: m = -l..l), from the Helgaker-Jorgensen-Olsen expansion."""
    carts = _cart_list(l)
    index = {c: k for k, c in enumerate(carts)}
    out = np.zeros((2 * l + 1, len(carts)))
    for row, m in enumerate(_sph_order(l)):
        am = abs(m)
        wm = 1 if m < 0 else 0
        norm = 1.0 / (2 ** am * math.factorial(l)) * math.sqrt(
            2.0 * math.factorial(l + am) * math.factorial(l - am) / (2.0 if m == 0 else 1.0))
        for t in range((l - am) // 2 + 1):
            for u in range(t + 1):
                for w in range(wm, am + 1, 2):
                    sign = -1.0 if (t + (w - wm) // 2) % 2 else 1.0
                    c = sign * 0.25 ** t * _binom(l, t) * _binom(l - t, am + t) * _binom(t, u) * _binom(am, w)
                    out[row, index[(2 * t + am - 2 * u - w, 2 * u + w, l - 2 * t - am)]] += norm * c
    return out


_SPH_CACHE: dict = {}


def _eval_cart(l, v):
    v = np.atleast_2d(v)
    return np.stack([v[:, 0] ** i * v[:, 1] ** j * v[:, 2] ** k for i, j, k in _cart_list(l)], axis=-1)


def solid_harmonics(l: int, vectors: np.ndarray) -> np.ndarray:
    """``S_lm(v)`` ``[npoints, 2l+1]``, every m scaled to unit mean square on the unit sphere."""
    if l not in _SPH_CACHE:
        t = cart_to_sph(l)
        # Gauss-Legendre in cos(theta) x uniform phi: exact for the degree-2l products.
        x, wx = np.polynomial.legendre.leggauss(l + 2)
        phi = np.linspace(0.0, 2 * np.pi, 2 * l + 3, endpoint=False)
        ct, ph = np.meshgrid(x, phi, indexing="ij")
        st = np.sqrt(1.0 - ct ** 2)
        pts = np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=-1).reshape(-1, 3)
        w = np.repeat(wx, len(phi)) * (2 * np.pi / len(phi))
        vals = _eval_cart(l, pts) @ t.T
        norms = np.sqrt((w[:, None] * vals ** 2).sum(0) / (4 * np.pi))
        _SPH_CACHE[l] = t / norms[:, None]
    return _eval_cart(l, vectors) @ _SPH_CACHE[l].T


_ROT_POINTS = None
_ROT_PINV: dict = {}


def rotation_matrices(rot: np.ndarray, lmax: int) -> list[np.ndarray]:
    """``D^l`` (l = 0..lmax) with ``S_l(rot^T v) = S_l(v) D^l`` for every v.

    A density ``sum_m c_m S_m(r)`` equals ``sum_m c'_m S_m(rot^T r)`` when
    ``c = D c'``. Solved by least squares on 64 fixed directions (exact: the
    harmonics of one l span a rotation-invariant space); the pseudo-inverse
    of the fixed directions' harmonics is computed once per l.
    """
    return [d[0] for d in rotation_matrices_many(np.asarray(rot)[None], lmax)]


def rotation_matrices_many(rots: np.ndarray, lmax: int) -> list[np.ndarray]:
    """:func:`rotation_matrices` of many rotations ``[k, 3, 3]`` at once: a list over l of ``[k, 2l+1, 2l+1]``."""
    global _ROT_POINTS
    if _ROT_POINTS is None:
        pts = np.random.default_rng(12345).standard_normal((64, 3))
        _ROT_POINTS = pts / np.linalg.norm(pts, axis=1, keepdims=True)
    k = rots.shape[0]
    rotated = np.einsum("pi,kij->kpj", _ROT_POINTS, rots).reshape(-1, 3)
    out = []
    for l in range(lmax + 1):
        if l not in _ROT_PINV:
            _ROT_PINV[l] = np.linalg.pinv(solid_harmonics(l, _ROT_POINTS))
        y = solid_harmonics(l, rotated).reshape(k, len(_ROT_POINTS), 2 * l + 1)
        out.append(np.einsum("mp,kpn->kmn", _ROT_PINV[l], y))
    return out


# ---------------------------------------------------------------- frames and descriptors

def water_frames(molecule: Molecule) -> np.ndarray:
    """Per atom, the rotation whose columns are its water's local axes in the laboratory frame.

    Atoms come as (O, H, H) per water, as ``water_cluster`` places them.
    """
    pos = molecule.positions
    nat = len(molecule.charges)
    if nat % 3 or any(tuple(molecule.charges[3 * k:3 * k + 3]) != (8, 1, 1) for k in range(nat // 3)):
        raise InputError("the learned guess is defined for water clusters in (O, H, H) order")
    w = np.asarray(pos, dtype=np.float64).reshape(nat // 3, 3, 3)  # [water, (O, H1, H2), xyz]
    z = (w[:, 1] - w[:, 0]) + (w[:, 2] - w[:, 0])
    z /= np.linalg.norm(z, axis=1, keepdims=True)
    x = w[:, 1] - w[:, 2]
    x -= np.sum(x * z, axis=1, keepdims=True) * z
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    rots = np.stack([x, np.cross(z, x), z], axis=2)  # columns: local axes
    return np.repeat(rots, 3, axis=0)


def _descriptor_rows(pos: np.ndarray, charges: np.ndarray, group: np.ndarray, rots: np.ndarray) -> np.ndarray:
    """Descriptors of atoms at ``pos`` (Bohr) whose neighbours are the atoms of the same ``group``
    (groups are contiguous runs of atoms)."""
    pos = pos / ANGSTROM_IN_BOHR  # Å
    centres = np.linspace(R_MIN, R_MAX, N_RADIAL)
    width = centres[1] - centres[0]
    nat = len(charges)
    n_ang = (L_DESCRIPTOR + 1) ** 2
    feats = np.zeros((nat, len(ELEMENTS), N_RADIAL * n_ang))
    # The (atom a, neighbour b) pairs within the cutoff, group by group; then every pair at once.
    pairs = []
    bounds = np.flatnonzero(np.diff(np.concatenate([[-1], group, [-2]]))) if nat else np.zeros(1, dtype=np.int64)
    for lo, hi in zip(bounds[:-1], bounds[1:]):
        r = np.linalg.norm(pos[None, lo:hi, :] - pos[lo:hi, None, :], axis=2)
        a, b = np.nonzero((r > 1e-8) & (r < R_MAX))
        pairs.append((a + lo, b + lo))
    ia = np.concatenate([p[0] for p in pairs]) if pairs else np.zeros(0, dtype=np.int64)
    ib = np.concatenate([p[1] for p in pairs]) if pairs else np.zeros(0, dtype=np.int64)
    if len(ia) == 0:
        return feats.reshape(nat, -1)
    d = pos[ib] - pos[ia]
    rp = np.linalg.norm(d, axis=1)
    local = np.einsum("pk,pkj->pj", d, rots[ia]) / rp[:, None]
    cut = 0.5 * (np.cos(np.pi * rp / R_MAX) + 1.0)
    radial = np.exp(-0.5 * ((rp[:, None] - centres[None, :]) / width) ** 2) * cut[:, None]
    ang = np.concatenate([solid_harmonics(l, local) for l in range(L_DESCRIPTOR + 1)], axis=1)
    outer = (radial[:, :, None] * ang[:, None, :]).reshape(len(ia), -1)
    # Pairs come ordered by atom a, so each atom's neighbours of one element are one run of rows.
    for e_i, z in enumerate(ELEMENTS):
        sel = np.flatnonzero(charges[ib] == z)
        if len(sel) == 0:
            continue
        atom = ia[sel]
        starts = np.flatnonzero(np.concatenate([[True], atom[1:] != atom[:-1]]))
        feats[atom[starts], e_i] = np.add.reduceat(outer[sel], starts, axis=0)
    return feats.reshape(nat, -1)


def descriptors(molecule: Molecule) -> np.ndarray:
    """``[natom, N_FEATURES]`` environment descriptors in each atom's water frame."""
    charges = np.asarray(molecule.charges)
    return _descriptor_rows(np.asarray(molecule.positions, dtype=np.float64), charges,
                            np.zeros(len(charges), dtype=np.int64), water_frames(molecule))


def descriptors_many(molecules) -> list[np.ndarray]:
    """:func:`descriptors` of many molecules, computed together (neighbours stay within a molecule)."""
    if not molecules:
        return []
    counts = [len(m.charges) for m in molecules]
    rows = _descriptor_rows(np.concatenate([np.asarray(m.positions, dtype=np.float64) for m in molecules]),
                            np.concatenate([np.asarray(m.charges) for m in molecules]),
                            np.repeat(np.arange(len(molecules)), counts),
                            np.concatenate([water_frames(m) for m in molecules]))
    return np.split(rows, np.cumsum(counts)[:-1])


# ---------------------------------------------------------------- coefficient layout

@functools.lru_cache(maxsize=None)
def shell_blocks(aux_name: str, z: int):
    """``([(start, l), ...], n)``: the (2l+1)-blocks of element z's auxiliary functions."""
    blocks = []
    at = 0
    for sh in load_basis(aux_name)[str(z)]:
        for _ in sh["coefficients"]:
            blocks.append((at, sh["l"]))
            at += 2 * sh["l"] + 1
    return blocks, at


@functools.lru_cache(maxsize=None)
def _lmax(aux_name: str) -> int:
    return max(sh["l"] for z in ELEMENTS for sh in load_basis(aux_name)[str(z)])


@functools.lru_cache(maxsize=None)
def _block_columns(aux_name: str, z: int) -> dict:
    """``{l: [n_blocks_l, 2l+1]}`` coefficient columns of element z's (2l+1)-blocks."""
    by_l: dict = {}
    for start, l in shell_blocks(aux_name, z)[0]:
        by_l.setdefault(l, []).append(np.arange(start, start + 2 * l + 1))
    return {l: np.asarray(v) for l, v in by_l.items()}


def _rotate_atoms(coeffs: np.ndarray, dms: list, columns: dict, inverse: bool = False) -> np.ndarray:
    """Rotate every (2l+1)-block of ``coeffs [atoms, n]`` by the atoms' ``D^l`` (``dms[l]`` ``[atoms, 2l+1, 2l+1]``)."""
    out = np.array(coeffs, dtype=np.float64, copy=True)
    for l, cols in columns.items():
        d = dms[l].transpose(0, 2, 1) if inverse else dms[l]
        out[:, cols] = np.einsum("aij,abj->abi", d, out[:, cols])
    return out


def atom_offsets(molecule: Molecule, aux_name: str) -> list[int]:
    offs, at = [], 0
    for z in molecule.charges:
        offs.append(at)
        at += shell_blocks(aux_name, z)[1]
    return offs


def to_local(molecule: Molecule, coeffs: np.ndarray, aux_name: str) -> dict:
    """Laboratory-frame coefficients -> ``{z: [atoms of z, n_coef_z]}`` in each atom's water frame.

    The local coefficients c' of an atom satisfy c = D(R) c' (``rotation_matrices``),
    so c' = D^T c (D is orthogonal).
    """
    offs = atom_offsets(molecule, aux_name)
    charges = np.asarray(molecule.charges)
    dms = _frame_rotations(molecule, aux_name)
    out = {}
    for z in sorted(set(molecule.charges)):
        atoms = np.nonzero(charges == z)[0]
        n = shell_blocks(aux_name, z)[1]
        lab = np.stack([coeffs[offs[a]:offs[a] + n] for a in atoms])
        out[z] = _rotate_atoms(lab, [d[atoms // 3] for d in dms], _block_columns(aux_name, z), inverse=True)
    return out


def _frame_rotations(molecule: Molecule, aux_name: str) -> list:
    """``rotation_matrices`` of every water's frame (its three atoms share it): a list over l of
    ``[waters, 2l+1, 2l+1]``."""
    rots = water_frames(molecule)
    return rotation_matrices_many(rots[0::3], _lmax(aux_name))


def to_lab(molecule: Molecule, local: dict, aux_name: str) -> np.ndarray:
    """Inverse of :func:`to_local`."""
    offs = atom_offsets(molecule, aux_name)
    charges = np.asarray(molecule.charges)
    dms = _frame_rotations(molecule, aux_name)
    total = offs[-1] + shell_blocks(aux_name, molecule.charges[-1])[1]
    out = np.zeros(total)
    for z in sorted(set(molecule.charges)):
        atoms = np.nonzero(charges == z)[0]
        n = shell_blocks(aux_name, z)[1]
        lab = _rotate_atoms(np.asarray(local[z], dtype=np.float64)[:len(atoms)], [d[atoms // 3] for d in dms],
                            _block_columns(aux_name, z))
        for k, a in enumerate(atoms):
            out[offs[a]:offs[a] + n] = lab[k]
    return out


def to_lab_many(molecules, locals_: list, aux_name: str) -> list[np.ndarray]:
    """:func:`to_lab` of many molecules: the frames of all waters rotated in one pass per element."""
    if not molecules:
        return []
    waters = [len(m.charges) // 3 for m in molecules]
    dms = rotation_matrices_many(np.concatenate([water_frames(m)[0::3] for m in molecules]), _lmax(aux_name))
    first_water = np.concatenate([[0], np.cumsum(waters)[:-1]])
    outs = []
    for m in molecules:
        offs = atom_offsets(m, aux_name)
        outs.append(np.zeros(offs[-1] + shell_blocks(aux_name, m.charges[-1])[1]))
    for z in ELEMENTS:
        n = shell_blocks(aux_name, z)[1]
        mol_of, atom_of, coef = [], [], []
        for k, (m, loc) in enumerate(zip(molecules, locals_)):
            atoms = np.nonzero(np.asarray(m.charges) == z)[0]
            if len(atoms) == 0:
                continue
            mol_of.append(np.full(len(atoms), k))
            atom_of.append(atoms)
            coef.append(np.asarray(loc[z], dtype=np.float64)[:len(atoms)])
        if not coef:
            continue
        mol_of, atom_of = np.concatenate(mol_of), np.concatenate(atom_of)
        water = first_water[mol_of] + atom_of // 3
        lab = _rotate_atoms(np.concatenate(coef), [d[water] for d in dms], _block_columns(aux_name, z))
        for k in np.unique(mol_of):
            sel = np.nonzero(mol_of == k)[0]
            offs = np.asarray(atom_offsets(molecules[k], aux_name))[atom_of[sel]]
            outs[k][(offs[:, None] + np.arange(n)[None, :]).reshape(-1)] = lab[sel].reshape(-1)
    return outs


def fitted_coefficients(metric: np.ndarray, three_center: np.ndarray, density_alpha) -> np.ndarray:
    """``d = M^-1 (P|uv) (2 D_alpha)``: the Coulomb-fitted coefficients of the total density."""
    naux = metric.shape[0]
    gamma = three_center.reshape(naux, -1) @ (2.0 * np.asarray(density_alpha)).reshape(-1)
    return np.linalg.solve(metric, gamma)


def coulomb_of_coefficients(three_center: np.ndarray, coeffs: np.ndarray) -> np.ndarray:
    """``J_uv = sum_P (P|uv) d_P``: the Coulomb matrix of a fitted density (= J[2 D_alpha] when d fits D)."""
    naux, n, _ = three_center.shape
    return (coeffs @ three_center.reshape(naux, n * n)).reshape(n, n)


def one_centre_metric(aux_name: str, z: int) -> np.ndarray:
    from hafnium_amd import integrals

    atom = Molecule((z,), np.zeros((1, 3)))
    return integrals.metric(BasisOnAtoms(aux_name, atom))


# ---------------------------------------------------------------- the model

@dataclass
class ElementModel:
    z: int
    feature_mean: np.ndarray
    feature_scale: np.ndarray
    target_mean: np.ndarray       # local coefficients
    whitening: np.ndarray         # U with G = U^T U (one-centre metric)
    target_scale: float
    params: np.ndarray            # flat MLP parameters (FP32 master)

    @property
    def n_coef(self) -> int:
        return len(self.target_mean)

    def whiten(self, local: np.ndarray) -> np.ndarray:
        return ((local - self.target_mean) @ self.whitening.T) / self.target_scale

    def unwhiten(self, y: np.ndarray) -> np.ndarray:
        return np.linalg.solve(self.whitening, (y * self.target_scale).T).T + self.target_mean

    def normalise(self, feats: np.ndarray) -> np.ndarray:
        return ((feats - self.feature_mean) / self.feature_scale).astype(np.float32)


class LearnedGuessModel:
    """Per-element networks and their normalisation; ``predict`` gives laboratory-frame coefficients."""

    def __init__(self, elements: dict, aux_name: str, *, device=None) -> None:
        self.elements = elements
        self.aux_name = aux_name
        self.device = device
        self._nets: dict = {}

    def net(self, z: int):
        from hafnium_amd.mlp import Mlp

        if z not in self._nets:
            em = self.elements[z]
            net = Mlp([N_FEATURES, *HIDDEN, em.n_coef], device=self.device)
            net.set_params(em.params)
            self._nets[z] = net
        return self._nets[z]

    def predict_local(self, molecule: Molecule) -> dict:
        torch = _torch()
        feats = descriptors(molecule)
        charges = np.asarray(molecule.charges)
        out = {}
        for z in ELEMENTS:
            sel = charges == z
            if not np.any(sel):
                continue
            em = self.elements[z]
            net = self.net(z)
            x = torch.from_numpy(em.normalise(feats[sel])).to(net.device)
            y = net.forward(x).cpu().numpy().astype(np.float64)
            out[z] = em.unwhiten(y)
        return out

    def predict(self, molecule: Molecule) -> np.ndarray:
        return to_lab(molecule, self.predict_local(molecule), self.aux_name)

    def predict_many(self, molecules) -> list[np.ndarray]:
        """Laboratory-frame coefficients of many molecules: one network pass per element for all atoms."""
        torch = _torch()
        feats = descriptors_many(molecules)
        charges = [np.asarray(m.charges) for m in molecules]
        local = [dict() for _ in molecules]
        for z in ELEMENTS:
            rows = [f[c == z] for f, c in zip(feats, charges)]
            total = sum(len(r) for r in rows)
            if total == 0:
                continue
            em = self.elements[z]
            net = self.net(z)
            x = torch.from_numpy(em.normalise(np.concatenate(rows))).to(net.device)
            y = em.unwhiten(net.forward(x).cpu().numpy().astype(np.float64))
            at = 0
            for k, r in enumerate(rows):
                local[k][z] = y[at:at + len(r)]
                at += len(r)
        return to_lab_many(molecules, local, self.aux_name)

    def save(self, path: str) -> None:
        arrays = {"aux_name": np.array(self.aux_name), "hidden": np.array(HIDDEN)}
        for z, em in self.elements.items():
            for key in ("feature_mean", "feature_scale", "target_mean", "whitening", "params"):
                arrays[f"{z}_{key}"] = getattr(em, key)
            arrays[f"{z}_target_scale"] = np.array(em.target_scale)
        np.savez_compressed(path, **arrays)

    @classmethod
    def load(cls, path: str = MODEL_PATH, *, device=None) -> "LearnedGuessModel":
        if not os.path.exists(path):
            raise InputError(f"no trained learned-guess model at {path}")
        with np.load(path, allow_pickle=False) as f:
            if tuple(int(h) for h in f["hidden"]) != HIDDEN:
                raise InputError("the stored model has another hidden width")
            elements = {}
            for z in ELEMENTS:
                elements[z] = ElementModel(z, f[f"{z}_feature_mean"], f[f"{z}_feature_scale"],
                                           f[f"{z}_target_mean"], f[f"{z}_whitening"],
                                           float(f[f"{z}_target_scale"]), f[f"{z}_params"])
            return cls(elements, str(f["aux_name"]), device=device)


# ---------------------------------------------------------------- data

def cluster_record(n_waters: int, seed: int, sad_cache: dict, *, device=None) -> dict:
    """Converge (H2O)_n in FP64 from the atomic guess; descriptors and local fitted coefficients per atom."""
    from hafnium_amd.clusters import atomic_density_guess, build_system, run_cluster

    system = build_system(water_cluster(n_waters, seed), device=device, keep_integrals=True)
    guess = atomic_density_guess(system.molecule, device=device, cache=sad_cache)
    result, _ = run_cluster(n_waters, seed, system=system, guess=guess)
    d = fitted_coefficients(system.metric, system.three_center, result.density.cpu().numpy())
    local = to_local(system.molecule, d, system.aux_basis)
    return {"features": descriptors(system.molecule), "charges": np.asarray(system.molecule.charges),
            "local": local, "energy": result.energy, "iterations": result.n_iterations}


def generate_dataset(seeds, sizes, path: str, *, device=None, log=print, checkpoint_every: int = 100,
                     keep_features: bool = True) -> dict:
    """Converged FP64 densities of ``water_cluster(sizes[i], seeds[i])`` as per-element arrays in ``path``
    (rewritten every ``checkpoint_every`` clusters, so an interrupted run keeps what it finished).

    Without ``keep_features`` only the targets and the (seed, size) of every cluster are stored;
    :func:`with_features` recomputes the descriptors from them (they depend on the geometry only).
    """
    sad_cache: dict = {}
    feats = {z: [] for z in ELEMENTS}
    targets = {z: [] for z in ELEMENTS}
    cluster_of = {z: [] for z in ELEMENTS}
    meta = []

    def save():
        arrays = {"meta": np.asarray(meta, dtype=np.float64)}
        for z in ELEMENTS:
            if keep_features:
                arrays[f"x_{z}"] = np.concatenate(feats[z])
            arrays[f"y_{z}"] = np.concatenate(targets[z])
            arrays[f"cluster_{z}"] = np.asarray(cluster_of[z], dtype=np.int32)
        np.savez_compressed(path, **arrays)
        return arrays

    t0 = time.perf_counter()
    for index, (seed, n) in enumerate(zip(seeds, sizes)):
        rec = cluster_record(int(n), int(seed), sad_cache, device=device)
        for z in ELEMENTS:
            sel = rec["charges"] == z
            feats[z].append(rec["features"][sel].astype(np.float32))
            targets[z].append(rec["local"][z])
            cluster_of[z].extend([index] * int(sel.sum()))
        meta.append((int(seed), int(n), rec["energy"], rec["iterations"]))
        log(f"[data] {index + 1}/{len(seeds)} n={n} seed={seed} E={rec['energy']:.10f} it={rec['iterations']} "
            f"t={time.perf_counter() - t0:.0f}s", flush=True)
        if checkpoint_every and (index + 1) % checkpoint_every == 0:
            save()
    return save()


def merge_datasets(parts: list[dict]) -> dict:
    """One dataset from several of :func:`generate_dataset` (cluster indices renumbered)."""
    if not parts:
        raise InputError("no dataset to merge")
    out = {"meta": np.concatenate([p["meta"] for p in parts])}
    for z in ELEMENTS:
        offsets = np.cumsum([0] + [int(p["meta"].shape[0]) for p in parts[:-1]])
        out[f"y_{z}"] = np.concatenate([p[f"y_{z}"] for p in parts])
        out[f"cluster_{z}"] = np.concatenate([p[f"cluster_{z}"] + o for p, o in zip(parts, offsets)]).astype(np.int32)
        if all(f"x_{z}" in p for p in parts):
            out[f"x_{z}"] = np.concatenate([p[f"x_{z}"] for p in parts])
    return out


def with_features(data: dict) -> dict:
    """``data`` with the per-element descriptors, recomputed from each cluster's (seed, size) in
    ``meta`` where the dataset does not store them (same order as the targets)."""
    if all(f"x_{z}" in data for z in ELEMENTS):
        return data
    feats = {z: [] for z in ELEMENTS}
    for seed, n in data["meta"][:, :2].astype(np.int64):
        mol = water_cluster(int(n), int(seed))
        x = descriptors(mol).astype(np.float32)
        charges = np.asarray(mol.charges)
        for z in ELEMENTS:
            feats[z].append(x[charges == z])
    out = dict(data)
    for z in ELEMENTS:
        out[f"x_{z}"] = np.concatenate(feats[z])
        if out[f"x_{z}"].shape[0] != out[f"y_{z}"].shape[0]:
            raise InputError("descriptors and targets do not line up")
    return out


# ---------------------------------------------------------------- training

def train(data: dict, *, aux_name: str = "def2-universal-jkfit", epochs: int = 60, batch: int = 256,
          lr: float = 1e-3, weight_decay: float = 1e-2, validation_fraction: float = 0.1, seed: int = 0,
          device=None, log=print) -> tuple["LearnedGuessModel", dict]:
    """Train one network per element on the GPU; returns the model with the best validation loss."""
    from hafnium_amd.mlp import Mlp

    torch = _torch()
    dev = torch.device(device if device is not None else "cuda:0")
    rng = np.random.default_rng(seed)
    n_clusters = int(data["meta"].shape[0])
    val_clusters = set(rng.choice(n_clusters, max(1, int(round(validation_fraction * n_clusters))),
                                  replace=False).tolist())
    elements = {}
    report = {}
    for z in ELEMENTS:
        x = np.asarray(data[f"x_{z}"], dtype=np.float64)
        local = np.asarray(data[f"y_{z}"], dtype=np.float64)
        is_val = np.isin(np.asarray(data[f"cluster_{z}"]), list(val_clusters))
        xt, lt = x[~is_val], local[~is_val]
        g = one_centre_metric(aux_name, z)
        u = np.linalg.cholesky(g).T
        fmean = xt.mean(0)
        fscale = np.maximum(xt.std(0), 1e-6)
        tmean = lt.mean(0)
        resid = (lt - tmean) @ u.T
        tscale = float(np.sqrt(np.mean(resid ** 2))) or 1.0
        em = ElementModel(z, fmean, fscale, tmean, u, tscale, np.zeros(0))
        net = Mlp([N_FEATURES, *HIDDEN, em.n_coef], seed=seed + z, device=dev)
        xs = torch.from_numpy(em.normalise(x)).to(dev)
        ys = torch.from_numpy(em.whiten(local).astype(np.float32)).to(dev)
        train_idx = np.nonzero(~is_val)[0].astype(np.int32)
        val_idx = torch.from_numpy(np.nonzero(is_val)[0].astype(np.int32)).to(dev)
        n_train = len(train_idx)
        if n_train == 0:
            raise InputError(f"no training atoms of element {z}")
        steps_per_epoch = (n_train + batch - 1) // batch  # the last minibatch may be short
        total = epochs * steps_per_epoch
        best = (math.inf, None, -1)
        step = 0
        t0 = time.perf_counter()
        for epoch in range(epochs):
            perm = torch.from_numpy(rng.permutation(train_idx)).to(dev)
            for s in range(steps_per_epoch):
                # Cosine decay to 5 % of the initial rate.
                rate = lr * (0.05 + 0.95 * 0.5 * (1.0 + math.cos(math.pi * step / total)))
                net.train_step(xs, ys, (perm, s * batch, min((s + 1) * batch, n_train)), lr=rate,
                               weight_decay=weight_decay)
                step += 1
            if epoch % 10 == 9 or epoch == epochs - 1:
                nv = int(val_idx.numel())
                if nv:
                    yv = net.forward(xs, (val_idx, 0, nv))
                    val = float(net.loss(yv, ys, (val_idx, 0, nv), scale=1.0 / nv).item())
                else:
                    val = math.nan
                if not val > best[0]:
                    best = (val, net.params.cpu().numpy().copy(), epoch)
                log(f"[train] z={z} epoch {epoch + 1}/{epochs} val={val:.4e} (per-atom whitened MSE x dim) "
                    f"t={time.perf_counter() - t0:.1f}s", flush=True)
        em.params = best[1]
        elements[z] = em
        # Baseline: the mean local density (what the network adds to).
        nv = int(is_val.sum())
        base = float(np.mean(np.sum(em.whiten(local[is_val]) ** 2, axis=1))) if nv else math.nan
        report[z] = {"train_atoms": int(n_train), "val_atoms": nv, "best_val": best[0], "best_epoch": best[2],
                     "mean_density_val": base, "seconds": time.perf_counter() - t0}
    return LearnedGuessModel(elements, aux_name, device=dev), report


# ---------------------------------------------------------------- superposition of water densities

WATER_OH_ANGSTROM = 0.96
WATER_ANGLE_DEG = 104.5


def reference_water() -> Molecule:
    """The cluster generator's rigid water in its own frame: O at the origin, the
    bisector along z, H1 - H2 along x (the frame :func:`water_frames` builds)."""
    half = math.radians(WATER_ANGLE_DEG) / 2.0
    s, c = WATER_OH_ANGSTROM * math.sin(half), WATER_OH_ANGSTROM * math.cos(half)
    return Molecule.from_angstrom(["O", "H", "H"], np.array([[0.0, 0.0, 0.0], [s, 0.0, c], [-s, 0.0, c]]))


def water_orbitals(*, basis: str = "cc-pvdz", aux: str = "def2-universal-jkfit", device=None,
                   cache: dict | None = None) -> np.ndarray:
    """Converged occupied orbitals ``[n_ao(H2O), 5]`` of the isolated reference water (FP64 SCF)."""
    from hafnium_amd.clusters import atomic_density_guess, build_system
    from hafnium_amd.scf import ScfOptions, solve_scf

    key = ("water-orbitals", basis, aux)
    if cache is not None and key in cache:
        return cache[key]
    mol = reference_water()
    system = build_system(mol, basis=basis, aux=aux, device=device)
    guess = atomic_density_guess(mol, basis=basis, device=device, cache=cache)
    result = solve_scf(system.hamiltonian, guess=guess,
                       options=ScfOptions(energy_tolerance=1e-11, gradient_tolerance=1e-8))
    c = result.mo_coefficients[:, :system.hamiltonian.n_alpha].cpu().numpy().copy()
    if cache is not None:
        cache[key] = c
    return c


@functools.lru_cache(maxsize=None)
def _orbital_rows(basis: str) -> list:
    """Per atom of a water (O, H, H): ``(row offset, {l: columns})`` of its orbital-basis blocks."""
    out, at = [], 0
    for z in (8, 1, 1):
        out.append((at, _block_columns(basis, z)))
        at += shell_blocks(basis, z)[1]
    return out


def molecular_factors(molecule: Molecule, c_ref: np.ndarray, basis: str = "cc-pvdz"):
    """Superposition of water densities: ``(D, C)`` with C block-diagonal over the waters, block k =
    the reference water's occupied orbitals rotated into water k's frame (c_lab = D(R) c_local per
    (2l+1)-block, as for the fitted coefficients), and D = C C^T."""
    rots = water_frames(molecule)[0::3]
    nw, (nw_ao, nocc) = len(rots), c_ref.shape
    lmax = max(max(cols) for _, cols in _orbital_rows(basis))
    dms = rotation_matrices_many(rots, lmax)
    blocks = np.repeat(c_ref[None], nw, axis=0)  # [waters, n_ao(H2O), nocc]
    for offset, cols in _orbital_rows(basis):
        for l, c in cols.items():
            rows = offset + c  # [n_blocks_l, 2l+1]
            blocks[:, rows, :] = np.einsum("wij,wbjo->wbio", dms[l], blocks[:, rows, :])
    factors = np.zeros((nw * nw_ao, nw * nocc))
    for k in range(nw):
        factors[k * nw_ao:(k + 1) * nw_ao, k * nocc:(k + 1) * nocc] = blocks[k]
    return factors @ factors.T, factors


# ---------------------------------------------------------------- the guess

@dataclass
class GuessStart:
    """A starting point for ``solve_scf``: density, its occupied orbitals and its Fock build."""

    density: object
    orbitals: object
    first_build: tuple  # (fock, terms) of hamiltonian.fock(density, orbitals)


def _aufbau_start(h, setup, fock_guess, precision):
    from hafnium_amd.scf import _density, _energy_terms

    _, c = setup.solve(fock_guess)
    density, orbitals = _density(h.dense, c, n_alpha=h.n_alpha)
    fock, terms = h.fock(density, orbitals, precision)
    return GuessStart(density, orbitals, (fock, terms)), _energy_terms(h, terms.cpu().numpy()).total


def atomic_iterate_start(system, sad_density: np.ndarray, *, precision: str = "fp64") -> GuessStart:
    """The fallback start alone: the aufbau density of the atomic guess's Fock matrix."""
    from hafnium_amd.scf import _Setup, ScfOptions

    torch = _torch()
    h = system.hamiltonian
    sad = torch.from_numpy(np.ascontiguousarray(sad_density)).to(h.device)
    fock, _ = h.fock(sad, None, precision)
    return _aufbau_start(h, _Setup(h, ScfOptions()), fock, precision)[0]


def learned_guess(system, model: LearnedGuessModel, sad_density: np.ndarray, *, precision: str = "fp64"):
    """Start for ``run_cluster`` from the predicted fitted density, with the atomic-guess fallback.

    F_sad = h + J[2 D_sad] - K[D_sad] is the Fock matrix of the atomic guess;
    F0 = h + J[d] - K[D_sad] uses the predicted coefficients d instead, with
    J[d] = sum_P q_P B[P], q = L^-1 (P|Q) d, on the GPU. The SCF starts from
    the aufbau density of F0 (the learned start, its Fock build handed on as
    the SCF's first iteration) unless its energy exceeds the energy of the
    atomic guess itself; then it starts from the atomic guess with F_sad as
    its first iteration, exactly as without the network (the fallback).
    Returns ``(GuessStart, info)``; ``info["used"]`` is "learned" or "sad".
    """
    from hafnium_amd.scf import _Setup, ScfOptions, _energy_terms

    torch = _torch()
    if system.metric is None:
        raise InputError("build the system with keep_integrals=True for the learned guess")
    h = system.hamiltonian
    tb = h.two_body
    dn = h.dense
    t0 = time.perf_counter()
    coeffs = torch.from_numpy(model.predict(system.molecule)).to(h.device).view(-1, 1).contiguous()
    metric = torch.as_tensor(system.metric, dtype=torch.float64).to(h.device).contiguous()
    charges = tb.fit_charges(dn.gemm(metric, coeffs)).view(-1)
    j_pred = tb.coulomb_from_charges(charges, precision=precision)
    sad = torch.from_numpy(np.ascontiguousarray(sad_density)).to(h.device)
    j_sad = tb.coulomb(sad, precision=precision)
    k_sad = tb.exchange(sad, precision=precision)
    f_sad = dn.combine(1.0, h.core, 2.0, j_sad, -1.0, k_sad)
    terms_sad = torch.empty(3, dtype=torch.float64, device=h.device)
    dn.dot(sad, h.core, terms_sad[0:1])
    dn.dot(sad, j_sad, terms_sad[1:2])
    dn.dot(sad, k_sad, terms_sad[2:3])
    f0 = dn.combine(1.0, h.core, 1.0, j_pred, -1.0, k_sad)
    learned, e_learned = _aufbau_start(h, _Setup(h, ScfOptions()), f0, precision)
    e_sad = _energy_terms(h, terms_sad.cpu().numpy()).total
    use = e_learned <= e_sad
    info = {"used": "learned" if use else "sad", "energy_learned": e_learned, "energy_sad": e_sad,
            "seconds": time.perf_counter() - t0}
    return (learned if use else GuessStart(sad, None, (f_sad, terms_sad))), info


def evaluate(model: "LearnedGuessModel", seeds, sizes, *, precision: str = "fp64", log=print) -> dict:
    """Held-out comparison, one cluster at a time: SCF from the atomic guess against SCF from the
    learned guess (same Hamiltonian, FP64 convergence to 1e-8 Eh). Times exclude integrals and the
    fit; the learned time includes the guess."""
    from hafnium_amd.clusters import atomic_density_guess, build_system, run_cluster

    torch = _torch()
    cache: dict = {}
    rows = []
    for seed, n in zip(seeds, sizes):
        system = build_system(water_cluster(int(n), int(seed)), keep_integrals=True)
        sad = atomic_density_guess(system.molecule, cache=cache)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ref, _ = run_cluster(int(n), int(seed), system=system, guess=sad, precision=precision)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        start, info = learned_guess(system, model, sad, precision=precision)
        got, _ = run_cluster(int(n), int(seed), system=system, guess=start, precision=precision)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        fb, _ = run_cluster(int(n), int(seed), system=system, guess=atomic_iterate_start(system, sad),
                            precision=precision)
        row = {"seed": int(seed), "n_waters": int(n), "sad_iterations": ref.n_iterations,
               "learned_iterations": got.n_iterations, "fallback_iterations": fb.n_iterations,
               "sad_seconds": t1 - t0, "learned_seconds": t2 - t1, "used": info["used"],
               "start_error_learned": info["energy_learned"] - ref.energy,
               "start_error_fallback": info["energy_sad"] - ref.energy,
               "energy_diff": max(abs(got.energy - ref.energy), abs(fb.energy - ref.energy))}
        rows.append(row)
        log(f"[eval] {len(rows)}/{len(seeds)} {row}", flush=True)
        del system
    summary = {"count": len(rows), "precision": precision,
               "mean_sad_iterations": float(np.mean([r["sad_iterations"] for r in rows])),
               "mean_learned_iterations": float(np.mean([r["learned_iterations"] for r in rows])),
               "mean_fallback_only_iterations": float(np.mean([r["fallback_iterations"] for r in rows])),
               "median_start_error_learned": float(np.median([r["start_error_learned"] for r in rows])),
               "median_start_error_fallback": float(np.median([r["start_error_fallback"] for r in rows])),
               "sad_seconds": float(sum(r["sad_seconds"] for r in rows)),
               "learned_seconds": float(sum(r["learned_seconds"] for r in rows)),
               "fallbacks": sum(1 for r in rows if r["used"] != "learned"),
               "max_energy_diff": float(max(r["energy_diff"] for r in rows)), "clusters": rows}
    return summary


# ---------------------------------------------------------------- command line

def _main(argv=None) -> int:
    p = argparse.ArgumentParser(prog="python -m hafnium_amd.learned_guess")
    sub = p.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate", help="converged FP64 cluster densities -> dataset .npz")
    g.add_argument("--count", type=int, default=400)
    g.add_argument("--sizes", default="2-8", help="water counts drawn uniformly from this range")
    g.add_argument("--seed0", type=int, default=100000)
    g.add_argument("--range", default=None, help="start:stop slice of the count clusters to converge here")
    g.add_argument("--no-features", action="store_true", help="store targets only (descriptors recomputed)")
    g.add_argument("--out", required=True)
    t = sub.add_parser("train", help="dataset .npz files -> model .npz")
    t.add_argument("--data", required=True, nargs="+")
    t.add_argument("--out", required=True)
    t.add_argument("--epochs", type=int, default=60)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--weight-decay", type=float, default=1e-2)
    t.add_argument("--batch", type=int, default=256)
    e = sub.add_parser("evaluate", help="held-out clusters: atomic vs learned guess -> JSON summary")
    e.add_argument("--model", default=MODEL_PATH)
    e.add_argument("--count", type=int, default=100)
    e.add_argument("--sizes", default="2-16")
    e.add_argument("--seed0", type=int, default=500000)
    e.add_argument("--precision", default="fp64")
    e.add_argument("--out", required=True)
    args = p.parse_args(argv)
    if args.cmd == "evaluate":
        import json

        lo, hi = (int(v) for v in args.sizes.split("-"))
        sizes = np.random.default_rng(args.seed0).integers(lo, hi + 1, args.count)
        seeds = args.seed0 + np.arange(args.count)
        summary = evaluate(LearnedGuessModel.load(args.model), seeds, sizes, precision=args.precision)
        with open(args.out, "w") as fh:
            json.dump(summary, fh, indent=1)
        print(json.dumps({k: v for k, v in summary.items() if k != "clusters"}), flush=True)
        return 0
    if args.cmd == "generate":
        lo, hi = (int(v) for v in args.sizes.split("-"))
        rng = np.random.default_rng(args.seed0)
        sizes = rng.integers(lo, hi + 1, args.count)
        seeds = args.seed0 + np.arange(args.count)
        start, stop = (0, args.count) if args.range is None else (int(v) for v in args.range.split(":"))
        generate_dataset(seeds[start:stop], sizes[start:stop], args.out, keep_features=not args.no_features)
    else:
        parts = []
        for path in args.data:
            with np.load(path, allow_pickle=False) as f:
                parts.append({k: f[k] for k in f.files})
        data = with_features(merge_datasets(parts))
        model, report = train(data, epochs=args.epochs, lr=args.lr, batch=args.batch,
                              weight_decay=args.weight_decay)
        model.save(args.out)
        for z, r in report.items():
            print(f"[train] element {z}: {r}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(_main())
