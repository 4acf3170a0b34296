"""Runs the host-emulated library (tests/emu/_out/emu_driver, AddressSanitizer)
on one molecule and compares every output with the CPU oracle -- test
infrastructure for checking kernel memory safety and numerics without a GPU.

Usage (from the repository root): python tests/emu/run_emu.py CASE
CASE: water | small_d | contracted_d | pure_small_d
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
from quax_amd import workloads  # noqa: E402
from quax_amd.basis import BasisSet, basis_transform  # noqa: E402


def case(name):
    from test_gpu_eri import contracted_d_cluster, small_d_cluster
    if name == "water":
        m = workloads.water_sto3g()
        return m, BasisSet.build(m, "sto-3g")
    if name == "small_d":
        return small_d_cluster()
    if name == "pure_small_d":
        m, b = small_d_cluster()
        return m, BasisSet(b.basis_name, b.shells, pure=True)
    if name == "contracted_d":
        return contracted_d_cluster()
    if name == "cluster2":
        m = workloads.cluster10()
        from quax_amd.molecule import Molecule
        m = Molecule([m.symbol(i) for i in range(2)], m.geometry()[:2])
        return m, BasisSet.build(m, "even-tempered")
    raise SystemExit(f"unknown case {name}")


def run(name, tol=1e-14, seed=3):
    mol, bs = case(name)
    nbf = bs.nbf()
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((nbf, nbf)) * 0.3
    D = X @ X.T / nbf
    l, atom, nprim, exps, coefs = bs.arrays()
    lines = [str(mol.natom()), " ".join(repr(float(v)) for v in mol.geometry().reshape(-1)),
             " ".join(repr(mol.charge(i)) for i in range(mol.natom())), str(len(l))]
    lines += [f"{l[s]} {atom[s]} {nprim[s]}" for s in range(len(l))]
    lines += [" ".join(repr(float(v)) for v in exps), " ".join(repr(float(v)) for v in coefs),
              f"{1 if bs.pure else 0} {tol!r}", " ".join(repr(float(v)) for v in D.reshape(-1))]
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.txt"), os.path.join(td, "out.bin")
        open(fin, "w").write("\n".join(lines) + "\n")
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
        r = subprocess.run([os.path.join(HERE, "_out", "emu_driver"), fin, fout], capture_output=True, text=True,
                           env=env, timeout=3600)
        if r.returncode != 0:
            print(r.stderr[-6000:])
            raise SystemExit(f"emu_driver failed with {r.returncode}")
        out = np.fromfile(fout)
    N = mol.natom()
    n2, n4 = nbf * nbf, nbf ** 4
    pos = 0
    res = {}
    if nbf <= 60:
        res["G"] = out[pos:pos + n4].reshape(nbf, nbf, nbf, nbf)
        pos += n4
    if nbf <= 40:
        res["dG"] = out[pos:pos + 3 * N * n4].reshape(3 * N, nbf, nbf, nbf, nbf)
        pos += 3 * N * n4
    for k, sz in (("J", n2), ("K", n2), ("g", 3 * N), ("S", n2), ("T", n2), ("V", n2), ("g1", 3 * N),
                  ("J32", n2), ("K32", n2), ("g32", 3 * N), ("Jmx", n2), ("Kmx", n2), ("gmx", 3 * N)):
        res[k] = out[pos:pos + sz]
        pos += sz
    assert pos == out.size
    # oracle (Cartesian) + transform
    ob = oracle.OracleBasis(mol, BasisSet(bs.basis_name, bs.shells, pure=False))
    Tm = basis_transform(bs)
    Gc = ob.compute_2e_int(tol)
    G = np.einsum("ip,jq,kr,ls,pqrs->ijkl", Tm, Tm, Tm, Tm, Gc, optimize=True)
    errs = {}
    if "G" in res:
        errs["G"] = np.abs(res["G"] - G).max()
    dGc = ob.eri_deriv_core1(tol)
    if "dG" in res:
        dG = np.einsum("ip,jq,kr,ls,xpqrs->xijkl", Tm, Tm, Tm, Tm, dGc, optimize=True)
        errs["dG"] = np.abs(res["dG"] - dG).max()
    errs["J"] = np.abs(res["J"].reshape(nbf, nbf) - np.einsum("pqrs,rs->pq", G, D)).max()
    errs["K"] = np.abs(res["K"].reshape(nbf, nbf) - np.einsum("prqs,rs->pq", G, D)).max()
    Dc = Tm.T @ D @ Tm
    g_ref = 2 * np.einsum("xpqrs,pq,rs->x", dGc, Dc, Dc, optimize=True) - np.einsum("xpqrs,pr,qs->x", dGc, Dc, Dc,
                                                                                       optimize=True)
    errs["g"] = np.abs(res["g"] - g_ref).max()
    S, T, V = (Tm @ x @ Tm.T for x in ob.oei())
    errs["S"] = np.abs(res["S"].reshape(nbf, nbf) - S).max()
    errs["T"] = np.abs(res["T"].reshape(nbf, nbf) - T).max()
    errs["V"] = np.abs(res["V"].reshape(nbf, nbf) - V).max()
    dS, dT, dV = ob.oei_deriv1()
    g1_ref = np.einsum("xpq,pq->x", dT + dV, Dc) + np.einsum("xpq,pq->x", dS, Dc)
    errs["g1"] = np.abs(res["g1"] - g1_ref).max()
    # mixed precision against the FP64 path of the same build
    for k in ("J", "K", "g"):
        errs[k + "32"] = np.abs(res[k + "32"] - res[k]).max()
        errs[k + "mx"] = np.abs(res[k + "mx"] - res[k]).max()
    return errs


if __name__ == "__main__":
    for nm in sys.argv[1:] or ["water"]:
        e = run(nm)
        print(nm, " ".join(f"{k}={v:.2e}" for k, v in e.items()), flush=True)
