"""CPU oracle for Quax's ERI path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this package, and only as the checker.  The product (``quax_amd``) never
imports it.

Contents:
* ctypes bindings to ``liboracle.so`` (qx_oracle.c, the C restatement of the
  reference's integral loops and of the HGP/Obara-Saika arithmetic Libint
  supplies);
* numpy restatements of the consumers: the index contract
  (utils.py:39-70, utils.h:131-172), nuclear repulsion (energy_utils.py:6-15),
  Cholesky orthogonalisation (energy_utils.py:29-36) and the RHF SCF loop
  (hartree_fock.py:8-119), plus the analytic RHF gradient assembled from
  integral derivatives.
"""
from __future__ import annotations

import ctypes
import itertools
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_boys.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
        L.or_basis_create.restype = ctypes.c_void_p
        L.or_basis_create.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, _ip, _ip, _ip, _dp, _dp]
        L.or_basis_destroy.argtypes = [ctypes.c_void_p]
        L.or_nbf.argtypes = [ctypes.c_void_p]
        L.or_get_coefs.argtypes = [ctypes.c_void_p, _dp]
        L.or_shell_quartet.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.or_shell_quartet_deriv1.argtypes = L.or_shell_quartet.argtypes
        L.or_shell_quartet_deriv2.argtypes = L.or_shell_quartet.argtypes
        L.or_schwarz.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp, _ip]
        L.or_schwarz.restype = ctypes.c_int
        L.or_compute_2e_int.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp]
        L.or_eri_deriv_core1.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp]
        L.or_compute_2e_deriv1.argtypes = [ctypes.c_void_p, ctypes.c_double, _ip, _dp]
        L.or_compute_2e_deriv1.restype = ctypes.c_int
        L.or_eri_deriv_core2.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp]
        L.or_compute_2e_deriv2.argtypes = [ctypes.c_void_p, ctypes.c_double, _ip, _dp]
        L.or_compute_2e_deriv2.restype = ctypes.c_int
        L.or_quartet_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, ctypes.c_int, _dp,
                                       ctypes.POINTER(ctypes.c_longlong)]
        L.or_quartet_batch.restype = ctypes.c_double
        L.or_quartet_batch_deriv1.argtypes = L.or_quartet_batch.argtypes
        L.or_quartet_batch_deriv1.restype = ctypes.c_double
        L.or_oei.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.or_oei_deriv1.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        _lib = L
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def boys(mmax: int, T: float) -> np.ndarray:
    out = np.zeros(mmax + 1)
    lib().or_boys(mmax, float(T), _p(out))
    return out


def ncart(l):
    return (l + 1) * (l + 2) // 2


class OracleBasis:
    """A molecule + Cartesian basis as the oracle sees it."""

    def __init__(self, molecule, basis):
        self.molecule = molecule
        self.basis = basis
        l, atom, nprim, exps, coefs = basis.arrays()
        self.l, self.atom, self.nprim = l, atom, nprim
        xyz = np.ascontiguousarray(molecule.geometry().reshape(-1))
        Zn = np.array([molecule.charge(i) for i in range(molecule.natom())], dtype=np.float64)
        self.natom = molecule.natom()
        self._keep = (xyz, Zn, l, atom, nprim, exps, coefs)
        self.h = lib().or_basis_create(self.natom, _p(xyz), _p(Zn), len(l), _p(l, _ip), _p(atom, _ip),
                                       _p(nprim, _ip), _p(exps), _p(coefs))
        self.nbf = lib().or_nbf(self.h)
        self.bf = np.concatenate([[0], np.cumsum([ncart(x) for x in l])]).astype(int)

    def __del__(self):
        try:
            lib().or_basis_destroy(self.h)
        except Exception:
            pass

    def normalized_coefs(self):
        out = np.zeros(len(self._keep[5]))
        lib().or_get_coefs(self.h, _p(out))
        return out

    # -- two-electron --
    def shell_quartet(self, a, b, c, d):
        n = [ncart(self.l[s]) for s in (a, b, c, d)]
        out = np.zeros(n)
        lib().or_shell_quartet(self.h, a, b, c, d, _p(out))
        return out

    def shell_quartet_deriv1(self, a, b, c, d):
        n = [ncart(self.l[s]) for s in (a, b, c, d)]
        out = np.zeros([12] + n)
        lib().or_shell_quartet_deriv1(self.h, a, b, c, d, _p(out))
        return out

    def shell_quartet_deriv2(self, a, b, c, d):
        """78 blocks: centre-coordinate pairs i <= j of 0..11, lexicographic."""
        n = [ncart(self.l[s]) for s in (a, b, c, d)]
        out = np.zeros([78] + n)
        lib().or_shell_quartet_deriv2(self.h, a, b, c, d, _p(out))
        return out

    def schwarz(self, tol):
        ns = len(self.l)
        vals = np.zeros((ns, ns))
        pairs = np.zeros(ns * (ns + 1), dtype=np.int32)
        npairs = lib().or_schwarz(self.h, tol, _p(vals), _p(pairs, _ip))
        return vals, pairs[:2 * npairs].reshape(-1, 2)

    def compute_2e_int(self, tol):
        n = self.nbf
        out = np.zeros((n, n, n, n))
        lib().or_compute_2e_int(self.h, tol, _p(out))
        return out

    def eri_deriv_core1(self, tol):
        n = self.nbf
        out = np.zeros((3 * self.natom, n, n, n, n))
        lib().or_eri_deriv_core1(self.h, tol, _p(out))
        return out

    def compute_2e_deriv1(self, tol, deriv_vec):
        n = self.nbf
        dv = np.ascontiguousarray(deriv_vec, dtype=np.int32)
        out = np.zeros((n, n, n, n))
        rc = lib().or_compute_2e_deriv1(self.h, tol, _p(dv, _ip), _p(out))
        if rc != 0:
            raise ValueError("oracle supports first-order deriv_vec only")
        return out

    def eri_deriv_core2(self, tol):
        n = self.nbf
        m = 3 * self.natom
        out = np.zeros((m * (m + 1) // 2, n, n, n, n))
        lib().or_eri_deriv_core2(self.h, tol, _p(out))
        return out

    def compute_2e_deriv2(self, tol, deriv_vec):
        n = self.nbf
        dv = np.ascontiguousarray(deriv_vec, dtype=np.int32)
        out = np.zeros((n, n, n, n))
        rc = lib().or_compute_2e_deriv2(self.h, tol, _p(dv, _ip), _p(out))
        if rc != 0:
            raise ValueError("compute_2e_deriv2 needs a second-order deriv_vec")
        return out

    def quartet_batch(self, quartets, nthreads=1, want=False, deriv=False):
        q = np.ascontiguousarray(quartets, dtype=np.int32).reshape(-1, 4)
        sizes = np.array([np.prod([ncart(self.l[s]) for s in row]) * (12 if deriv else 1) for row in q],
                         dtype=np.int64)
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        out = np.zeros(int(off[-1])) if want else None
        fn = lib().or_quartet_batch_deriv1 if deriv else lib().or_quartet_batch
        tot = fn(self.h, len(q), _p(q, _ip), int(nthreads),
                 _p(out) if want else ctypes.cast(None, _dp),
                 off.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
        return (tot, out, off) if want else tot

    # -- one-electron --
    def oei(self):
        n = self.nbf
        S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
        lib().or_oei(self.h, _p(S), _p(T), _p(V))
        return S, T, V

    def oei_deriv1(self):
        n = self.nbf
        shp = (3 * self.natom, n, n)
        dS, dT, dV = np.zeros(shp), np.zeros(shp), np.zeros(shp)
        lib().or_oei_deriv1(self.h, _p(dS), _p(dT), _p(dV))
        return dS, dT, dV


# ---------------------------------------------------------------- index contract
def how_many_derivs(k, n):
    """utils.py:39-47: C(3k+n-1, n)."""
    return math.comb(3 * k + n - 1, n)


def generate_multi_index_lookup(nparams, deriv_order):
    """utils.h:142-172: combinations with replacement, lexicographic."""
    return [list(c) for c in itertools.combinations_with_replacement(range(nparams), deriv_order)]


def get_deriv_vec_idx(deriv_vec):
    """utils.py:49-70."""
    dv = np.asarray(deriv_vec, dtype=int)
    order = int(dv.sum())
    for idx, c in enumerate(itertools.combinations_with_replacement(range(dv.shape[0]), order)):
        t = np.zeros_like(dv)
        for i in c:
            t[i] += 1
        if np.array_equal(t, dv):
            return idx
    raise ValueError("deriv_vec not found")


# ---------------------------------------------------------------- RHF
def nuclear_repulsion(geom, Z):
    """energy_utils.py:6-15."""
    g = np.asarray(geom).reshape(-1, 3)
    e = 0.0
    for i in range(len(Z)):
        for j in range(i):
            e += Z[i] * Z[j] / np.linalg.norm(g[i] - g[j])
    return e


def nuclear_repulsion_grad(geom, Z):
    g = np.asarray(geom).reshape(-1, 3)
    out = np.zeros_like(g)
    for i in range(len(Z)):
        for j in range(len(Z)):
            if i == j:
                continue
            r = g[i] - g[j]
            out[i] -= Z[i] * Z[j] * r / np.linalg.norm(r) ** 3
    return out.reshape(-1)


def rhf(S, T, V, G, nelec, enuc, maxit=100, spectral_shift=False, conv=1e-12, damping=False,
        damp_factor=0.5, jk=None):
    """hartree_fock.py:8-119 restated in numpy (guess D = 0, Cholesky
    orthogonalisation, optional 1e-9 spectral shift, optional damping for the
    first 10 iterations, convergence on |dE| and RMS of the orbital gradient).
    ``jk(D) -> (J, K)`` may replace the tensor contraction."""
    n = S.shape[0]
    ndocc = nelec // 2
    A = np.linalg.inv(np.linalg.cholesky(S)).T
    H = T + V
    shift = np.diag(np.linspace(0, 1, n) * 1e-9) if spectral_shift else np.zeros((n, n))

    def jkf(D):
        if jk is not None:
            return jk(D)
        J = np.einsum("pqrs,rs->pq", G, D, optimize=True)
        K = np.einsum("prqs,rs->pq", G, D, optimize=True)
        return J, K

    def it(F, D):
        E = np.einsum("pq,pq->", F + H, D) + enuc
        Fp = A.T @ F @ A + shift
        eps, C2 = np.linalg.eigh(Fp)
        C = A @ C2
        Cocc = C[:, :ndocc]
        return E, Cocc @ Cocc.T, C, eps

    D = np.zeros((n, n))
    J, K = jkf(D)
    F = H + 2 * J - K
    D_old = D.copy()          # the loop carry starts from the zero guess (hartree_fock.py:97-99)
    E, D, C, eps = it(F, D)
    for k in range(maxit):
        if damping and k < 10:
            D = D_old * damp_factor + D * damp_factor
        D_old = D.copy()
        J, K = jkf(D)
        F = H + 2 * J - K
        # orbital gradient of the density that built F.  (The reference
        # evaluates it with the NEW density, which commutes with F by
        # construction, so only |dE| gates its convergence.)
        err = A @ (F @ D @ S - S @ D @ F) @ A
        E_new, D, C, eps = it(F, D)
        drms = np.sqrt(np.mean(err ** 2))
        dE = E - E_new
        E = E_new
        if abs(dE) <= conv and drms <= conv:
            break
    return dict(E=E, D=D, C=C, eps=eps, F=F, ndocc=ndocc, iterations=k + 1)


def rhf_gradient(scf, dS, dT, dV, dG, geom, Z):
    """Analytic RHF gradient consistent with the reference's energy
    E = sum (F+H) D + E_nuc (hartree_fock.py:59):
    dE = 2 D.dH + sum (2 D_pq D_rs - D_pr D_qs) dG_pqrs - 2 W.dS + dE_nuc."""
    D, C, eps, nd = scf["D"], scf["C"], scf["eps"], scf["ndocc"]
    W = C[:, :nd] @ np.diag(eps[:nd]) @ C[:, :nd].T
    g = 2 * np.einsum("xpq,pq->x", dT + dV, D)
    g -= 2 * np.einsum("xpq,pq->x", dS, W)
    g += 2 * np.einsum("xpqrs,pq,rs->x", dG, D, D, optimize=True)
    g -= np.einsum("xpqrs,pr,qs->x", dG, D, D, optimize=True)
    g += nuclear_repulsion_grad(geom, Z)
    return g
