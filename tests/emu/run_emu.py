"""Runs the host-emulated library (tests/emu/_out/emu_driver, AddressSanitizer)
on one molecule and compares every output with the CPU oracle -- test
infrastructure for checking kernel memory safety and numerics without a GPU.

Usage (from the repository root): python tests/emu/run_emu.py CASE
CASE: water | small_d | contracted_d | pure_small_d | cluster2 | fragments
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


def _driver(mol, bs, D, tol, env_extra=None):
    l, atom, nprim, exps, coefs = bs.arrays()
    lines = [str(mol.natom()), " ".join(repr(float(v)) for v in mol.geometry().reshape(-1)),
             " ".join(repr(mol.charge(i)) for i in range(mol.natom())), str(len(l))]
    lines += [f"{l[s]} {atom[s]} {nprim[s]}" for s in range(len(l))]
    lines += [" ".join(repr(float(v)) for v in exps), " ".join(repr(float(v)) for v in coefs),
              f"{1 if bs.pure else 0} {tol!r}", " ".join(repr(float(v)) for v in D.reshape(-1))]
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.txt"), os.path.join(td, "out.bin")
        open(fin, "w").write("\n".join(lines) + "\n")
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", **(env_extra or {}))
        r = subprocess.run([os.path.join(HERE, "_out", "emu_driver"), fin, fout], capture_output=True, text=True,
                           env=env, timeout=3600)
        if r.returncode != 0:
            print(r.stderr[-6000:])
            raise SystemExit(f"emu_driver failed with {r.returncode}")
        return np.fromfile(fout)


def run_fragments(tol=1e-14, seed=5):
    """Water, H2 and CO (STO-3G) as three fragments of one context: J, K and
    the two-electron gradient of a block-diagonal density against the oracle
    on each molecule alone."""
    from quax_amd.basis import Shell
    from quax_amd.molecule import Molecule
    mols = [workloads.water_sto3g(), Molecule(["H", "H"], [[0, 0, 0], [0, 0, 1.4]]),
            Molecule(["C", "O"], [[0, 0, 0], [0, 0, 2.13]])]
    bases = [BasisSet.build(m, "sto-3g") for m in mols]
    symbols, coords, shells, frag = [], [], [], []
    aoff = 0
    for k, (m, b) in enumerate(zip(mols, bases)):
        symbols += [m.symbol(i) for i in range(m.natom())]
        coords.append(np.asarray(m.geometry()).reshape(-1, 3))
        shells += [Shell(s.l, s.atom + aoff, list(s.exps), list(s.coefs)) for s in b.shells]
        frag += [k] * m.natom()
        aoff += m.natom()
    from quax_amd.molecule import Molecule as M
    comb = M(symbols, np.vstack(coords))
    cb = BasisSet("sto-3g", shells)
    nb = [b.nbf() for b in bases]
    off = np.concatenate([[0], np.cumsum(nb)]).astype(int)
    rng = np.random.default_rng(seed)
    D = np.zeros((off[-1], off[-1]))
    Ds = []
    for k, n in enumerate(nb):
        X = rng.standard_normal((n, n)) * 0.3
        Ds.append(X @ X.T / n)
        D[off[k]:off[k + 1], off[k]:off[k + 1]] = Ds[-1]
    out = _driver(comb, cb, D, tol, {"EMU_FRAGMENTS": " ".join(map(str, frag))})
    n, N = off[-1], comb.natom()
    n2, n4 = n * n, n ** 4
    pos = n4 if n <= 60 else 0          # G (not compared: no cross-fragment integrals)
    pos += 3 * N * n4 if n <= 40 else 0
    J = out[pos:pos + n2].reshape(n, n); pos += n2
    K = out[pos:pos + n2].reshape(n, n); pos += n2
    g = out[pos:pos + 3 * N]
    errs = {"J": 0.0, "K": 0.0, "g": 0.0, "Jx": 0.0}
    nat = np.concatenate([[0], np.cumsum([m.natom() for m in mols])]).astype(int)
    for k, (m, b) in enumerate(zip(mols, bases)):
        ob = oracle.OracleBasis(m, b)
        G = ob.compute_2e_int(tol)
        s_ = slice(off[k], off[k + 1])
        errs["J"] = max(errs["J"], np.abs(J[s_, s_] - np.einsum("pqrs,rs->pq", G, Ds[k])).max())
        errs["K"] = max(errs["K"], np.abs(K[s_, s_] - np.einsum("prqs,rs->pq", G, Ds[k])).max())
        dG = ob.eri_deriv_core1(tol)
        gk = 2 * np.einsum("xpqrs,pq,rs->x", dG, Ds[k], Ds[k]) - np.einsum("xpqrs,pr,qs->x", dG, Ds[k], Ds[k])
        errs["g"] = max(errs["g"], np.abs(g[3 * nat[k]:3 * nat[k + 1]] - gk).max())
        other = np.ones(n, dtype=bool)
        other[s_] = False
        errs["Jx"] = max(errs["Jx"], np.abs(J[s_][:, other]).max(), np.abs(K[s_][:, other]).max())
    return errs


if __name__ == "__main__":
    if sys.argv[1:] == ["fragments"]:
        e = run_fragments()
        print("fragments", " ".join(f"{k}={v:.2e}" for k, v in e.items()), flush=True)
        sys.exit(0)
    for nm in sys.argv[1:] or ["water"]:
        e = run(nm)
        print(nm, " ".join(f"{k}={v:.2e}" for k, v in e.items()), flush=True)
