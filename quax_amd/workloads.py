"""Seeded synthetic molecules of the benchmark and parity workloads.

* water_sto3g: RHF/STO-3G water, O (0,0,0), H (0,0,1), H (0,1,0) bohr.
* cluster10: 6 C + 4 H, uniform in an 8-bohr cube, minimum separation
  1.8 bohr (seed 2), even-tempered basis, spherical (280 functions) or
  Cartesian (300).
* cluster10_subset4: the first four atoms of cluster10 (4 C, 24 electrons,
  112 spherical / 120 Cartesian functions).
* cluster24: 14 C + 10 H, same rule (seed 4), 672 spherical functions.
* batch256: 256 molecules, 3-12 atoms each, elements C/H/O drawn with
  probabilities 0.4/0.4/0.2; a draw with an odd electron count is redrawn
  (a neutral singlet needs an even count).
"""
from __future__ import annotations

import numpy as np

from .molecule import Molecule, synthetic_cluster, water_config1

SEED_CLUSTER10 = 2
SEED_CLUSTER24 = 4
SEED_BATCH = 5


def water_sto3g() -> Molecule:
    return water_config1()


def cluster10() -> Molecule:
    return synthetic_cluster(6, 4, seed=SEED_CLUSTER10)


def cluster10_subset4() -> Molecule:
    m = cluster10()
    return Molecule([m.symbol(i) for i in range(4)], m.geometry()[:4])


def cluster24() -> Molecule:
    return synthetic_cluster(14, 10, seed=SEED_CLUSTER24)


def _place(symbols, rng, box=8.0, min_sep=1.8):
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


def batch256(count: int = 256):
    rng = np.random.default_rng(SEED_BATCH)
    zmap = {"C": 6, "H": 1, "O": 8}
    out = []
    while len(out) < count:
        natom = int(rng.integers(3, 13))
        symbols = list(rng.choice(["C", "H", "O"], size=natom, p=[0.4, 0.4, 0.2]))
        if sum(zmap[s] for s in symbols) % 2:
            continue
        out.append(_place(symbols, rng))
    return out
