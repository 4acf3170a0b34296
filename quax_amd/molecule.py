"""Psi4-free molecule stand-in and the xyz file boundary of the ERI path.

The reference passes geometry to the native module through a file: Psi4 writes
``geom.xyz`` in Angstrom (core.py:318-320) and Libint reads it back with
``read_dotxyz`` (utils.h:4-9, libint_interface.cc:41), converting to bohr.
``Molecule`` duck-types the Psi4 calls the path makes (``geometry()``,
``natom()``, ``charge(i)``, ``molecular_charge()``, ``multiplicity()``,
``save_xyz_file()``) and ``read_xyz`` is the Libint-side reader.  Both sides
use the same bohr/Angstrom factor (CODATA 2018), so the round trip is exact to
the 17 significant digits written.
"""
from __future__ import annotations

import numpy as np

from .basis import SYMBOL_TO_Z, Z_TO_SYMBOL

BOHR_TO_ANGSTROM = 0.529177210903


class Molecule:
    def __init__(self, symbols, coords_bohr, charge: int = 0, multiplicity: int = 1):
        self._symbols = [s.capitalize() for s in symbols]
        self._xyz = np.asarray(coords_bohr, dtype=np.float64).reshape(-1, 3).copy()
        if len(self._symbols) != self._xyz.shape[0]:
            raise ValueError("symbols and coordinates differ in length")
        for s in self._symbols:
            if s not in SYMBOL_TO_Z:
                raise ValueError(f"unknown element '{s}'")
        self._charge = int(charge)
        self._mult = int(multiplicity)

    # -- Psi4 Molecule surface used by quax.core (core.py:315-333) --
    def geometry(self) -> np.ndarray:
        return self._xyz.copy()

    def natom(self) -> int:
        return len(self._symbols)

    def symbol(self, i: int) -> str:
        return self._symbols[i]

    def Z(self, i: int) -> int:
        return SYMBOL_TO_Z[self._symbols[i]]

    def charge(self, i: int) -> float:
        """Nuclear charge of atom i (psi4 ``Molecule.charge``)."""
        return float(SYMBOL_TO_Z[self._symbols[i]])

    def molecular_charge(self) -> int:
        return self._charge

    def multiplicity(self) -> int:
        return self._mult

    def nelectron(self) -> int:
        return int(sum(self.Z(i) for i in range(self.natom()))) - self._charge

    def set_geometry(self, coords_bohr) -> None:
        self._xyz = np.asarray(coords_bohr, dtype=np.float64).reshape(-1, 3).copy()

    def save_xyz_file(self, path: str, save_in_angstrom: bool = True) -> None:
        f = BOHR_TO_ANGSTROM if save_in_angstrom else 1.0
        with open(path, "w") as fh:
            fh.write(f"{self.natom()}\n\n")
            for s, r in zip(self._symbols, self._xyz):
                fh.write(f"{s} {r[0] * f:.17g} {r[1] * f:.17g} {r[2] * f:.17g}\n")

    @classmethod
    def from_xyz(cls, path: str, charge: int = 0, multiplicity: int = 1) -> "Molecule":
        symbols, coords = read_xyz(path)
        return cls(symbols, coords, charge, multiplicity)

    def __repr__(self):
        return f"Molecule(natom={self.natom()}, charge={self._charge}, mult={self._mult})"


def read_xyz(path: str):
    """Libint ``read_dotxyz`` semantics: first line atom count, second a
    comment, then ``symbol x y z`` in Angstrom; returns bohr."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    if not lines:
        raise ValueError(f"{path}: empty xyz file")
    try:
        n = int(lines[0].split()[0])
    except (ValueError, IndexError):
        raise ValueError(f"{path}: first line must hold the atom count")
    if len(lines) < 2 + n:
        raise ValueError(f"{path}: expected {n} atom lines")
    symbols, coords = [], []
    for ln in lines[2:2 + n]:
        parts = ln.split()
        if len(parts) < 4:
            raise ValueError(f"{path}: malformed atom line '{ln}'")
        sym = parts[0]
        if sym.isdigit():
            sym = Z_TO_SYMBOL[int(sym)]
        symbols.append(sym.capitalize())
        coords.append([float(v) / BOHR_TO_ANGSTROM for v in parts[1:4]])
    return symbols, np.asarray(coords, dtype=np.float64)


def water_config1() -> Molecule:
    """RHF/STO-3G water of the plain CPU-runnable case: O (0,0,0), H (0,0,1),
    H (0,1,0) bohr (the README example geometry)."""
    return Molecule(["O", "H", "H"], [[0, 0, 0], [0, 0, 1], [0, 1, 0]])


def water_reference_test() -> Molecule:
    """The water geometry of the reference's own tests (test_gradients.py:9-15), bohr."""
    return Molecule(["O", "H", "H"], [[-0.000007070942, 0.125146536460, 0.0],
                                      [-1.424097055410, -0.993053750648, 0.0],
                                      [1.424209276385, -0.993112599269, 0.0]])


def synthetic_cluster(n_c: int, n_h: int, seed: int, box: float = 8.0, min_sep: float = 1.8,
                      n_o: int = 0) -> Molecule:
    """Seeded synthetic cluster: positions uniform in a ``box``-bohr cube,
    rejection-sampled to a minimum pair separation; atoms ordered C, O, H."""
    rng = np.random.default_rng(seed)
    symbols = ["C"] * n_c + ["O"] * n_o + ["H"] * n_h
    pts = []
    tries = 0
    while len(pts) < len(symbols):
        tries += 1
        if tries > 1_000_000:
            raise RuntimeError("rejection sampling did not converge")
        p = rng.uniform(0.0, box, size=3)
        if all(np.linalg.norm(p - q) >= min_sep for q in pts):
            pts.append(p)
    return Molecule(symbols, np.asarray(pts))
