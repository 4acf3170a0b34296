"""Basis-set library and the Psi4-free basis/molecule stand-ins the ERI path needs.

The reference builds its basis twice: Psi4 for ``nbf()``/``name()``
(core.py:327) and Libint from the named library file
(libint_interface.cc:48-70, shells forced Cartesian by ``set_pure(false)``).
Both read named files that ship with those packages.  Here the two named sets
the path uses are built in:

* ``sto-3g`` for H, C, N, O (Hehre-Stewart-Pople, as distributed by the Basis
  Set Exchange); ``SP`` shells are split into an s shell followed by a p shell,
  as Libint's g94 reader does.
* ``even-tempered``: per atom, s exponents 0.1*3^k (k=0..5), p 0.15*3^k
  (k=0..3), d 0.4*3^k (k=0..1), each shell one primitive; oxygen exponents are
  scaled by 1.5.  This is the synthetic-cluster basis.

Shells are returned with raw (unnormalised) contraction coefficients; the
native library normalises them the way Libint does (x^l normalisation of each
primitive, then unit self-overlap of the contraction).

Cartesian component order within a shell is the Libint/CCA "standard" order
(xx, xy, xz, yy, yz, zz for d).  In spherical mode (``pure=True``) shells with
l >= 2 become 2l+1 real solid harmonics ordered m = -l..l; s and p shells are
unchanged (x, y, z).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

SYMBOL_TO_Z = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10}
Z_TO_SYMBOL = {v: k for k, v in SYMBOL_TO_Z.items()}

# Exponents/coefficients: STO-3G (1s and 2sp expansions scaled by the Slater
# exponents of each element).
_STO3G_1S_C = [0.15432897, 0.53532814, 0.44463454]
_STO3G_2S_C = [-0.09996723, 0.39951283, 0.70011547]
_STO3G_2P_C = [0.15591627, 0.60768372, 0.39195739]
_STO3G = {
    "H": [(0, [3.42525091, 0.62391373, 0.16885540], _STO3G_1S_C)],
    "C": [(0, [71.6168370, 13.0450960, 3.5305122], _STO3G_1S_C),
          (0, [2.9412494, 0.6834831, 0.2222899], _STO3G_2S_C),
          (1, [2.9412494, 0.6834831, 0.2222899], _STO3G_2P_C)],
    "N": [(0, [99.1061690, 18.0523120, 4.8856602], _STO3G_1S_C),
          (0, [3.7804559, 0.8784966, 0.2857144], _STO3G_2S_C),
          (1, [3.7804559, 0.8784966, 0.2857144], _STO3G_2P_C)],
    "O": [(0, [130.7093200, 23.8088610, 6.4436083], _STO3G_1S_C),
          (0, [5.0331513, 1.1695961, 0.3803890], _STO3G_2S_C),
          (1, [5.0331513, 1.1695961, 0.3803890], _STO3G_2P_C)],
}


def _even_tempered(symbol: str):
    scale = 1.5 if symbol == "O" else 1.0
    shells = []
    for k in range(6):
        shells.append((0, [scale * 0.1 * 3.0 ** k], [1.0]))
    for k in range(4):
        shells.append((1, [scale * 0.15 * 3.0 ** k], [1.0]))
    for k in range(2):
        shells.append((2, [scale * 0.4 * 3.0 ** k], [1.0]))
    return shells


BASIS_NAMES = ("sto-3g", "even-tempered")


def element_shells(basis_name: str, symbol: str):
    name = basis_name.lower()
    if name == "sto-3g":
        if symbol not in _STO3G:
            raise ValueError(f"sto-3g is not available for element {symbol}")
        return _STO3G[symbol]
    if name in ("even-tempered", "et"):
        return _even_tempered(symbol)
    raise ValueError(f"unknown basis set '{basis_name}' (known: {', '.join(BASIS_NAMES)})")


def ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def nfunc(l: int, pure: bool) -> int:
    return 2 * l + 1 if (pure and l >= 2) else ncart(l)


def cart_components(l: int):
    """(lx, ly, lz) in Libint/CCA standard order."""
    out = []
    for i in range(l + 1):
        lx = l - i
        for j in range(i + 1):
            out.append((lx, i - j, j))
    return out


@dataclass
class Shell:
    l: int
    atom: int
    exps: List[float]
    coefs: List[float]


@dataclass
class BasisSet:
    """Duck-types the part of ``psi4.core.BasisSet`` the path calls
    (``nbf()``, ``name()``; tei.py:20-23, ints.py:19) and carries the shells
    Libint would have built (libint_interface.cc:48-83)."""
    basis_name: str
    shells: List[Shell]
    pure: bool = False
    _shell2bf: List[int] = field(default_factory=list)

    def __post_init__(self):
        off = 0
        self._shell2bf = []
        for s in self.shells:
            self._shell2bf.append(off)
            off += nfunc(s.l, self.pure)
        self._nbf = off

    @classmethod
    def build(cls, molecule, basis_name: str, pure: bool = False) -> "BasisSet":
        shells = []
        for a in range(molecule.natom()):
            for (l, exps, coefs) in element_shells(basis_name, molecule.symbol(a)):
                shells.append(Shell(l, a, list(exps), list(coefs)))
        return cls(basis_name, shells, pure)

    def nbf(self) -> int:
        return self._nbf

    def ncart(self) -> int:
        return sum(ncart(s.l) for s in self.shells)

    def name(self) -> str:
        return self.basis_name.upper()

    def nshell(self) -> int:
        return len(self.shells)

    def shell2bf(self):
        return list(self._shell2bf)

    def shell2atom(self):
        return [s.atom for s in self.shells]

    def max_l(self) -> int:
        return max((s.l for s in self.shells), default=0)

    def arrays(self):
        """Flat arrays in the layout the C ABI takes."""
        l = np.array([s.l for s in self.shells], dtype=np.int32)
        atom = np.array([s.atom for s in self.shells], dtype=np.int32)
        nprim = np.array([len(s.exps) for s in self.shells], dtype=np.int32)
        exps = np.array([e for s in self.shells for e in s.exps], dtype=np.float64)
        coefs = np.array([c for s in self.shells for c in s.coefs], dtype=np.float64)
        return l, atom, nprim, exps, coefs


# ---------------------------------------------------------------- spherical
def _solid_harmonic_coefs(l: int) -> np.ndarray:
    """Rows m=-l..l, columns Cartesian components (standard order): monomial
    coefficients of the real solid harmonics r^l S_lm, unnormalised.  Uses the
    closed-form expansion (Helgaker, Jorgensen & Olsen, Molecular
    Electronic-Structure Theory, eq. 6.4.47):
    sum_{t,u,v} (-1)^(t+v-v_m) (1/4)^t C(l,t) C(l-t,|m|+t) C(t,u) C(|m|,2v)
    x^(2t+|m|-2(u+v)) y^(2(u+v)) z^(l-2t-|m|), v_m = 0 (m>=0) or 1/2 (m<0)."""
    comps = cart_components(l)
    index = {c: i for i, c in enumerate(comps)}
    M = np.zeros((2 * l + 1, len(comps)))
    for mi, m in enumerate(range(-l, l + 1)):
        am = abs(m)
        for t in range((l - am) // 2 + 1):
            for u in range(t + 1):
                # 2v runs over even values for m >= 0 and odd values for m < 0
                twov_values = range(0, am + 1, 2) if m >= 0 else range(1, am + 1, 2)
                for twov in twov_values:
                    sign = (-1) ** (t + (twov - (0 if m >= 0 else 1)) // 2)
                    c = sign * 0.25 ** t * math.comb(l, t) * math.comb(l - t, am + t) \
                        * math.comb(t, u) * math.comb(am, twov)
                    py = 2 * u + twov
                    px = 2 * t + am - py
                    pz = l - 2 * t - am
                    if px < 0 or pz < 0:
                        continue
                    M[mi, index[(px, py, pz)]] += c
    return M


def _df(n: int) -> float:
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


def cart_overlap_shared_norm(l: int) -> np.ndarray:
    """Overlap between the Cartesian components of one shell when all share the
    x^l normalisation (origin-centred, same exponent): entries
    (2(lx+lx')-1)!!... / (2l-1)!! when all summed powers are even."""
    comps = cart_components(l)
    S = np.zeros((len(comps), len(comps)))
    for i, a in enumerate(comps):
        for j, b in enumerate(comps):
            tot = [a[k] + b[k] for k in range(3)]
            if any(t % 2 for t in tot):
                continue
            S[i, j] = np.prod([_df(t - 1) for t in tot]) / _df(2 * l - 1)
    return S


def sph_transform(l: int) -> np.ndarray:
    """(2l+1, ncart(l)) Cartesian -> spherical transform with unit-norm rows
    under the shared-normalisation Cartesian metric."""
    if l < 2:
        return np.eye(ncart(l))
    M = _solid_harmonic_coefs(l)
    S = cart_overlap_shared_norm(l)
    norms = np.sqrt(np.einsum("mi,ij,mj->m", M, S, M))
    return M / norms[:, None]


def basis_transform(basis: BasisSet) -> np.ndarray:
    """Block-diagonal (nbf, ncart) matrix taking Cartesian AO quantities to the
    basis' own functions (identity for Cartesian bases)."""
    nc = basis.ncart()
    T = np.zeros((basis.nbf(), nc))
    r = c = 0
    for s in basis.shells:
        blk = sph_transform(s.l) if basis.pure else np.eye(ncart(s.l))
        T[r:r + blk.shape[0], c:c + blk.shape[1]] = blk
        r += blk.shape[0]
        c += blk.shape[1]
    return T
