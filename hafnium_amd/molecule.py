"""Molecules, basis sets on atoms, and the seeded water-cluster generator.

Atomic units inside (CODATA 2018 Bohr, the reference's
hafnium/molecule/units.py:6-9); positions are given in Ångström by default
and converted, as the reference's ``Molecule`` does
(hafnium/molecule/molecule.py:21-30). The nuclear repulsion follows
hafnium/molecule/molecule.py:45-56.

Basis data (BSE JSON, H and O of sto-3g, cc-pvdz, def2-svp and
def2-universal-jkfit) lives in ``hafnium_amd/data/basis_h_o.json``.
"""

from __future__ import annotations

import ctypes
import json
import math
import os
from dataclasses import dataclass

import numpy as np

from hafnium_amd.errors import InputError

BOHR_IN_ANGSTROM = 0.529177210903
ANGSTROM_IN_BOHR = 1.0 / BOHR_IN_ANGSTROM
SYMBOLS = {"H": 1, "O": 8}
DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "basis_h_o.json")


@dataclass(frozen=True)
class Molecule:
    """Nuclear charges and positions (Bohr)."""

    charges: tuple
    positions: np.ndarray  # [natom, 3] Bohr

    @classmethod
    def from_angstrom(cls, symbols, positions_angstrom) -> "Molecule":
        pos = np.asarray(positions_angstrom, dtype=np.float64) * ANGSTROM_IN_BOHR
        try:
            charges = tuple(SYMBOLS[s] for s in symbols)
        except KeyError as error:
            raise InputError(f"element {error.args[0]} has no bundled basis data") from None
        return cls(charges, pos)

    @property
    def n_electrons(self) -> int:
        return int(sum(self.charges))

    def nuclear_repulsion(self) -> float:
        total = 0.0
        for a in range(len(self.charges)):
            for b in range(a):
                total += self.charges[a] * self.charges[b] / float(np.linalg.norm(self.positions[a] - self.positions[b]))
        return total


def load_basis(name: str) -> dict:
    with open(DATA) as fh:
        sets = json.load(fh)["basis_sets"]
    try:
        return sets[name.lower()]
    except KeyError:
        raise InputError(f"basis {name!r} is not bundled (have {sorted(sets)})") from None


class _Basis(ctypes.Structure):
    _fields_ = [("nshell", ctypes.c_int), ("l", ctypes.c_void_p), ("nprim", ctypes.c_void_p),
                ("ncontr", ctypes.c_void_p), ("centers", ctypes.c_void_p), ("exps", ctypes.c_void_p),
                ("coefs", ctypes.c_void_p)]


class BasisOnAtoms:
    """The shells of a basis set placed on a molecule's atoms (C-ABI ``hfjk_basis``)."""

    def __init__(self, name: str, molecule: Molecule) -> None:
        data = load_basis(name)
        ls, nprim, ncontr, centers, exps, coefs = [], [], [], [], [], []
        for z, pos in zip(molecule.charges, molecule.positions):
            for sh in data[str(z)]:
                ls.append(sh["l"])
                nprim.append(len(sh["exponents"]))
                ncontr.append(len(sh["coefficients"]))
                centers.extend(float(x) for x in pos)
                exps.extend(sh["exponents"])
                for row in sh["coefficients"]:
                    coefs.extend(row)
        self.name = name
        self._l = np.asarray(ls, dtype=np.int32)
        self._nprim = np.asarray(nprim, dtype=np.int32)
        self._ncontr = np.asarray(ncontr, dtype=np.int32)
        self._centers = np.asarray(centers, dtype=np.float64)
        self._exps = np.asarray(exps, dtype=np.float64)
        self._coefs = np.asarray(coefs, dtype=np.float64)
        self.n_functions = int(sum((2 * l + 1) * c for l, c in zip(ls, ncontr)))

    def c_struct(self) -> _Basis:
        ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        return _Basis(len(self._l), ptr(self._l), ptr(self._nprim), ptr(self._ncontr), ptr(self._centers),
                      ptr(self._exps), ptr(self._coefs))


def water_cluster(n_waters: int, seed: int, *, oh: float = 0.96, angle_deg: float = 104.5,
                  min_oo: float = 2.6, radius_factor: float = 2.5, max_tries: int = 100000) -> Molecule:
    """A seeded (H2O)_n: rigid waters (O-H 0.96 Å, H-O-H 104.5°), random orientations,
    oxygens rejection-sampled in a sphere of radius 2.5 n^(1/3) Å with O-O >= 2.6 Å."""
    rng = np.random.default_rng(seed)
    radius = radius_factor * n_waters ** (1.0 / 3.0)
    half = math.radians(angle_deg) / 2.0
    local_h = np.array([[oh * math.sin(half), 0.0, oh * math.cos(half)],
                        [-oh * math.sin(half), 0.0, oh * math.cos(half)]])
    oxygens: list[np.ndarray] = []
    tries = 0
    while len(oxygens) < n_waters:
        tries += 1
        if tries > max_tries:
            raise InputError(f"could not place {n_waters} waters in the sphere")
        p = rng.uniform(-radius, radius, 3)
        if np.linalg.norm(p) > radius:
            continue
        if all(np.linalg.norm(p - o) >= min_oo for o in oxygens):
            oxygens.append(p)
    symbols, positions = [], []
    for o in oxygens:
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        rot = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        symbols += ["O", "H", "H"]
        positions += [o, o + rot @ local_h[0], o + rot @ local_h[1]]
    return Molecule.from_angstrom(symbols, np.asarray(positions))
