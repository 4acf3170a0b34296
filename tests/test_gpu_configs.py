"""The benchmark molecules at full size on the GPU.

* cluster10 (6 C + 4 H, 300 Cartesian functions, Schwarz 1e-12): all 8-fold
  unique integrals materialised on the device (6.15 GB); a seeded sample of
  shell quartets from every class list against the oracle at 1e-12 per
  integral; the direct J/K kernels against J and K contracted from those
  stored integrals over all 8 index permutations; the RHF energy at the
  converged density from both; the adjoint two-electron gradient against
  central finite differences.
* cluster10_subset4 (first four atoms, 112 spherical functions): the analytic
  RHF gradient (adjoint kernel for the two-electron part) against central
  finite differences of the RHF energy, step 1e-4 bohr.
* cluster24 (14 C + 10 H, 672 spherical functions): the RHF energy and
  gradient with FP32 recursion for quartets below a Schwarz-bound threshold
  against the all-FP64 path, |dE| <= 1e-8 Eh and |dg| <= 1e-7 Eh/bohr.
"""
import math

import numpy as np
import pytest

import oracle
from quax_amd import workloads
from quax_amd.basis import BasisSet
from quax_amd.molecule import Molecule

pytestmark = pytest.mark.gpu


def _ncart(l):
    return (l + 1) * (l + 2) // 2


def _density(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n)) * 0.3
    return X @ X.T / n


def _jk_from_store(torch, store, lists, pairs, bf_first, D, chunk=1 << 22):
    """J_pq = sum_rs (pq|rs) D_rs and K_pq = sum_rs (pr|qs) D_rs from the
    unique integrals: every stored element is expanded into its 8 index
    permutations, each weighted 1/2 per coincidence (a = b, c = d, ab = cd of
    the canonical shell quartet), which counts each element of the full
    tensor once."""
    dev = store.device
    n = D.shape[0]
    Dt = torch.as_tensor(D, device=dev).reshape(-1)
    J = torch.zeros(n * n, dtype=torch.float64, device=dev)
    K = torch.zeros(n * n, dtype=torch.float64, device=dev)
    pairs_t = torch.as_tensor(pairs, device=dev, dtype=torch.int64)
    bf_t = torch.as_tensor(bf_first, device=dev, dtype=torch.int64)
    for L in lists:
        la, lb, lc, ld = L["l4"]
        na, nb, nc, nd = (_ncart(x) for x in L["l4"])
        N, nq = na * nb * nc * nd, L["nquart"]
        comp = torch.arange(N, device=dev)
        ci, cj = comp // (nb * nc * nd), (comp // (nc * nd)) % nb
        ck, cl = (comp // nd) % nc, comp % nd
        p1all = torch.as_tensor(L["p1"], device=dev, dtype=torch.int64)
        p2all = torch.as_tensor(L["p2"], device=dev, dtype=torch.int64)
        step = max(1, chunk // N)
        for t0 in range(0, nq, step):
            t1 = min(nq, t0 + step)
            p1, p2 = p1all[t0:t1], p2all[t0:t1]
            sa, sb = pairs_t[p1, 0], pairs_t[p1, 1]
            sc, sd = pairs_t[p2, 0], pairs_t[p2, 1]
            f = torch.ones(t1 - t0, dtype=torch.float64, device=dev)
            f = torch.where(sa == sb, 0.5 * f, f)
            f = torch.where(sc == sd, 0.5 * f, f)
            f = torch.where(p1 == p2, 0.5 * f, f)
            I = (bf_t[sa][None, :] + ci[:, None]).reshape(-1)
            Jx = (bf_t[sb][None, :] + cj[:, None]).reshape(-1)
            Kx = (bf_t[sc][None, :] + ck[:, None]).reshape(-1)
            Lx = (bf_t[sd][None, :] + cl[:, None]).reshape(-1)
            idx = L["base"] + comp[:, None] * nq + torch.arange(t0, t1, device=dev)[None, :]
            v = (store[idx.reshape(-1)].reshape(N, -1) * f[None, :]).reshape(-1)
            for (p, q, r, s) in ((I, Jx, Kx, Lx), (Jx, I, Kx, Lx), (I, Jx, Lx, Kx), (Jx, I, Lx, Kx),
                                 (Kx, Lx, I, Jx), (Lx, Kx, I, Jx), (Kx, Lx, Jx, I), (Lx, Kx, Jx, I)):
                J.index_add_(0, p * n + q, v * Dt[r * n + s])
                K.index_add_(0, p * n + r, v * Dt[q * n + s])
    return J.reshape(n, n), K.reshape(n, n)


def _e2(torch, eng, D):
    J, K = eng.jk(D)
    Dt = torch.as_tensor(D, device=J.device)
    return float(torch.sum(Dt * (2 * J - K)))


def test_cluster10_full_size(gpu):
    import torch
    from quax_amd.engine import EriEngine
    from quax_amd.methods.hartree_fock import restricted_hartree_fock
    mol = workloads.cluster10()
    bs = BasisSet.build(mol, "even-tempered")            # Cartesian: 300 functions
    ob = oracle.OracleBasis(mol, bs)
    bf_first = np.concatenate([[0], np.cumsum([_ncart(s.l) for s in bs.shells])]).astype(np.int64)
    with EriEngine(mol, bs, ints_tol=1e-12) as eng:
        assert eng.nbf == 300
        nq_total = eng.num_quartets()
        store = eng.store()
        lists, pairs = eng.store_map()
        assert len(lists) == 21
        assert sum(L["nquart"] for L in lists) == nq_total   # single-primitive shells
        # -- per-integral parity on a seeded sample of every class list
        rng = np.random.default_rng(11)
        checked, worst = 0, 0.0
        for L in lists:
            nq = L["nquart"]
            N = math.prod(_ncart(x) for x in L["l4"])
            take = np.sort(rng.choice(nq, size=min(nq, 110), replace=False))
            idx = L["base"] + np.arange(N)[:, None] * nq + take[None, :]
            vals = store[torch.as_tensor(idx.reshape(-1), device=store.device)].cpu().numpy().reshape(N, -1)
            for j, t in enumerate(take):
                a, b = pairs[L["p1"][t]]
                c, d = pairs[L["p2"][t]]
                ref = ob.shell_quartet(int(a), int(b), int(c), int(d)).reshape(-1)
                worst = max(worst, float(np.abs(vals[:, j] - ref).max()))
                checked += 1
        assert checked >= 2000
        assert worst < 1e-12, worst
        # -- converged RHF density (direct J/K), then J/K and E from the store
        E, aux = restricted_hartree_fock(eng, mol.nelectron(), {"spectral_shift": False, "diis": True,
                                                                "conv": 1e-10, "maxit": 200},
                                         return_aux_data=True)
        D = aux["D"].cpu().numpy()
        Jd, Kd = (x.cpu().numpy() for x in eng.jk(D))
        Js, Ks = (x.cpu().numpy() for x in _jk_from_store(torch, store, lists, pairs, bf_first, D))
        del store
        torch.cuda.empty_cache()
        assert np.abs(Jd - Js).max() < 1e-11 * max(1.0, np.abs(Js).max())
        assert np.abs(Kd - Ks).max() < 1e-11 * max(1.0, np.abs(Ks).max())
        H = aux["H"].cpu().numpy()
        Ed = float(np.sum((2 * H + 2 * Jd - Kd) * D)) + aux["Enuc"]
        Es = float(np.sum((2 * H + 2 * Js - Ks) * D)) + aux["Enuc"]
        assert abs(Ed - Es) < 1e-10
        assert abs(Ed - E) < 1e-8           # the SCF energy at its own convergence
        g = eng.eri_grad(D).cpu().numpy()
        e2_0 = _e2(torch, eng, D)
    assert abs(g.reshape(-1, 3).sum(0)).max() < 1e-9   # translational invariance
    # -- adjoint gradient vs central differences of E2 at fixed D (step 1e-4)
    h = 1e-4
    xyz0 = np.asarray(mol.geometry(), dtype=np.float64).reshape(-1)
    for k in (0, 13, 29):
        e = []
        for s in (+1, -1):
            x = xyz0.copy()
            x[k] += s * h
            m = Molecule([mol.symbol(i) for i in range(mol.natom())], x.reshape(-1, 3))
            with EriEngine(m, bs, ints_tol=1e-12) as e1:
                e.append(_e2(torch, e1, D))
        fd = (e[0] - e[1]) / (2 * h)
        assert abs(g[k] - fd) < 1e-6 * max(1.0, abs(fd)), (k, g[k], fd)
    assert np.isfinite(e2_0)


def test_subset4_rhf_gradient_vs_finite_differences(gpu):
    from quax_amd.engine import EriEngine
    from quax_amd.methods.hartree_fock import restricted_hartree_fock, rhf_gradient
    mol = workloads.cluster10_subset4()
    bs = BasisSet.build(mol, "even-tempered", pure=True)
    assert bs.nbf() == 112
    opts = {"spectral_shift": False, "diis": True, "conv": 1e-12, "maxit": 300}

    def energy(xyz, want_grad=False):
        m = Molecule([mol.symbol(i) for i in range(mol.natom())], xyz.reshape(-1, 3))
        with EriEngine(m, bs, ints_tol=1e-14) as eng:
            E, aux = restricted_hartree_fock(eng, m.nelectron(), dict(opts), return_aux_data=True)
            assert aux["drms"] <= 1e-12
            return (E, rhf_gradient(eng, aux)) if want_grad else E

    xyz0 = np.asarray(mol.geometry(), dtype=np.float64).reshape(-1)
    E0, g = energy(xyz0, True)
    h = 1e-4
    fd = np.zeros_like(g)
    for k in range(xyz0.size):
        xp, xm = xyz0.copy(), xyz0.copy()
        xp[k] += h
        xm[k] -= h
        fd[k] = (energy(xp) - energy(xm)) / (2 * h)
    assert np.abs(g - fd).max() < 1e-7, np.abs(g - fd).max()
    assert np.abs(g.reshape(-1, 3).sum(0)).max() < 1e-9


def test_cluster24_mixed_precision_vs_fp64(gpu):
    from quax_amd.engine import EriEngine
    from quax_amd.methods.hartree_fock import restricted_hartree_fock, rhf_gradient
    mol = workloads.cluster24()
    bs = BasisSet.build(mol, "even-tempered", pure=True)
    assert bs.nbf() == 672
    opts = {"spectral_shift": False, "diis": True, "conv": 1e-10, "maxit": 300}
    with EriEngine(mol, bs, ints_tol=1e-12) as eng:
        nq = eng.num_quartets()
        E64, aux64 = restricted_hartree_fock(eng, mol.nelectron(), dict(opts), return_aux_data=True)
        g64 = rhf_gradient(eng, aux64)
        # the headline threshold from a cold start, the others warm-started
        # from the FP64 density
        for thr, warm in ((1e-8, False), (1e-10, True), (1e-6, True)):
            n64, n32 = eng.mixed_quartets(thr)
            assert n64 + n32 == nq and n32 > 0
            o = dict(opts, precision="mixed", fp32_threshold=thr)
            Em, auxm = restricted_hartree_fock(eng, mol.nelectron(), o, return_aux_data=True,
                                               D0=aux64["D"] if warm else None)
            gm = rhf_gradient(eng, auxm, precision="mixed", fp32_threshold=thr)
            print(f"cluster24 mixed @ {thr:g}: {n32 / nq:.1%} quartets in FP32, dE {Em - E64:.2e}, "
                  f"max dg {np.abs(gm - g64).max():.2e}", flush=True)
            assert abs(Em - E64) <= 1e-8
            assert np.abs(gm - g64).max() <= 1e-7
