"""Writes the committed golden fixtures under tests/golden/ (run on the CPU).

water_sto3g.npz   RHF/STO-3G water (O (0,0,0), H (0,0,1), H (0,1,0) bohr):
                  full ERI tensor (7^4), eri_deriv_core(1) tensor (9 x 7^4),
                  S, T, V, HF energy and analytic gradient, from the CPU oracle
                  (itself pinned against an independent McMurchie-Davidson
                  code, closed forms and finite differences in tests/).
geometries.json   the seeded synthetic molecules of the cluster/batch
                  workloads, so a change of generator is caught.

Usage: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from quax_amd import workloads  # noqa: E402
from quax_amd.basis import BasisSet  # noqa: E402

TOL = 1e-14


def water():
    mol = workloads.water_sto3g()
    bs = BasisSet.build(mol, "sto-3g")
    ob = oracle.OracleBasis(mol, bs)
    G = ob.compute_2e_int(TOL)
    dG = ob.eri_deriv_core1(TOL)
    S, T, V = ob.oei()
    Z = np.array([mol.charge(i) for i in range(mol.natom())])
    enuc = oracle.nuclear_repulsion(mol.geometry(), Z)
    scf = oracle.rhf(S, T, V, G, mol.nelectron(), enuc, conv=1e-13, maxit=200)
    dS, dT, dV = ob.oei_deriv1()
    g = oracle.rhf_gradient(scf, dS, dT, dV, dG, mol.geometry(), Z)
    np.savez_compressed(os.path.join(HERE, "water_sto3g.npz"), eri=G, eri_deriv1=dG, S=S, T=T, V=V,
                        energy=np.array(scf["E"]), gradient=g, geometry=mol.geometry())
    print("water: E =", scf["E"], "iterations", scf["iterations"])


def geometries():
    out = {}
    for name, mol in [("cluster10", workloads.cluster10()), ("cluster24", workloads.cluster24())]:
        out[name] = {"symbols": [mol.symbol(i) for i in range(mol.natom())], "xyz_bohr": mol.geometry().tolist()}
    b = workloads.batch256()
    out["batch256"] = [{"symbols": [m.symbol(i) for i in range(m.natom())], "xyz_bohr": m.geometry().tolist()}
                       for m in b]
    with open(os.path.join(HERE, "geometries.json"), "w") as fh:
        json.dump(out, fh, indent=0)


if __name__ == "__main__":
    water()
    geometries()
