"""Batched RHF energies and gradients: many molecules of uneven size share
one integral context and one J/K (and one adjoint-gradient) launch per
angular-momentum class per SCF iteration.

The molecules become fragments of one context (qx_initialize_fragments):
no shell pair or quartet spans two molecules, so J, K and the ERI gradient
of the block-diagonal batch density are the per-molecule results side by
side.  One-electron integrals, their gradient and the nuclear terms come
from one small context per molecule (they couple every nucleus).  The SCF
runs in lock step: one batched J/K build, then per molecule the same
update as restricted_hartree_fock (hartree_fock.py:25-99; DIIS on the
orbital gradient of the density that built F), with every molecule's
orthogonalised Fock matrix diagonalised in one padded batched host eigh.
Converged molecules keep their density while the others iterate.
"""
from __future__ import annotations

import numpy as np

from .._lib import QX_PREC_FP64, QX_PREC_MIXED
from ..basis import BasisSet, Shell
from ..molecule import Molecule
from .hartree_fock import nuclear_repulsion, nuclear_repulsion_grad

_PAD = 1.0e4  # diagonal of the padding rows: above every occupied orbital energy


def batched_rhf(molecules, basis_name="even-tempered", pure=True, ints_tol=1e-12, conv=1e-10, maxit=200,
                diis_space=20, precision="fp64", fp32_threshold=1e-8, gradient=True, device=0, graph=False,
                purify_iters=100):
    """Returns (energies, gradients or None, aux) for a list of neutral
    closed-shell molecules; aux holds iterations per molecule.  graph: the
    iterations of all molecules replayed from one captured HIP graph
    (GraphSCF: the single-molecule iteration of restricted_hartree_fock per
    molecule, DIIS on, no spectral shift, densities by purification on the
    device, converged molecules skipped; a molecule whose DIIS error stalls at
    the purification's rounding floor finishes in the host-driven loop);
    otherwise the host-driven lock-step loop (_eager_scf)."""
    import time
    import torch
    from ..engine import EriEngine
    t_start = time.perf_counter()
    dev = torch.device("cuda", device)
    dt = torch.float64
    prec = QX_PREC_MIXED if precision == "mixed" else QX_PREC_FP64
    bases = [BasisSet.build(m, basis_name, pure=pure) for m in molecules]
    # one context for the whole batch, molecules as fragments
    symbols, coords, shells, frag = [], [], [], []
    aoff = 0
    for k, (m, b) in enumerate(zip(molecules, bases)):
        symbols += [m.symbol(i) for i in range(m.natom())]
        coords.append(np.asarray(m.geometry()).reshape(-1, 3))
        shells += [Shell(s.l, s.atom + aoff, list(s.exps), list(s.coefs)) for s in b.shells]
        frag += [k] * m.natom()
        aoff += m.natom()
    comb = Molecule(symbols, np.vstack(coords))
    eng = EriEngine(comb, BasisSet(basis_name, shells, pure=pure), ints_tol=ints_tol, device=device, fragments=frag)
    nb = [b.nbf() for b in bases]
    off = np.concatenate([[0], np.cumsum(nb)]).astype(int)
    nat = [m.natom() for m in molecules]
    aof = np.concatenate([[0], np.cumsum(nat)]).astype(int)
    B, nmax, ntot = len(molecules), max(nb), int(off[-1])
    # per-molecule one-electron parts
    one = [EriEngine(m, b, ints_tol=ints_tol, device=device) for m, b in zip(molecules, bases)]
    H, S, A, ndocc, Enuc = [], [], [], [], []
    for m, e1 in zip(molecules, one):
        s_, t_, v_ = e1.oei()
        H.append(t_ + v_)
        S.append(s_)
        A.append(torch.linalg.inv(torch.linalg.cholesky(s_)).T.contiguous())
        ndocc.append(m.nelectron() // 2)
        Z = [m.charge(i) for i in range(m.natom())]
        Enuc.append(nuclear_repulsion(m.geometry(), Z))
    blk = [slice(int(off[k]), int(off[k + 1])) for k in range(B)]
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    timing = {"setup_s": t0 - t_start}
    if graph:
        E, D, C_all, eps_all, iters, done, stalled = _graph_scf(eng, H, S, A, ndocc, Enuc, off, blk, ntot, conv,
                                                                maxit, diis_space, prec, fp32_threshold,
                                                                purify_iters, dev)
        timing.update(graph_s=time.perf_counter() - t0, stalled=int(stalled.sum()))
        if stalled.any():
            # the last iterations of molecules whose purified densities hit
            # their rounding floor: eigh-based, from the graph's densities
            E2, D2, C2, eps2, it2, done2 = _eager_scf(eng, H, S, A, ndocc, Enuc, blk, nb, ntot, conv, maxit,
                                                      diis_space, prec, fp32_threshold, dev,
                                                      D0=[D[b_, b_] for b_ in blk], active=stalled)
            for k in np.flatnonzero(stalled):
                E[k], C_all[k], eps_all[k] = E2[k], C2[k], eps2[k]
                D[blk[k], blk[k]] = D2[blk[k], blk[k]]
                iters[k] += it2[k]
                done[k] = done2[k]
        timing["scf_s"] = time.perf_counter() - t0
        return _finish(eng, one, molecules, D, C_all, eps_all, blk, aof, ndocc, E, iters, done, nb, gradient, prec,
                       fp32_threshold, timing)
    E, D, C_all, eps_all, iters, done = _eager_scf(eng, H, S, A, ndocc, Enuc, blk, nb, ntot, conv, maxit, diis_space,
                                                   prec, fp32_threshold, dev)
    timing["scf_s"] = time.perf_counter() - t0
    return _finish(eng, one, molecules, D, C_all, eps_all, blk, aof, ndocc, E, iters, done, nb, gradient, prec,
                   fp32_threshold, timing)


def _eager_scf(eng, H, S, A, ndocc, Enuc, blk, nb, ntot, conv, maxit, diis_space, prec, fp32_threshold, dev,
               D0=None, active=None):
    """The host-driven lock-step loop over the molecules flagged in `active`
    (all by default), each from its block of D0 (zero by default); the other
    molecules' blocks stay as D0 gives them."""
    import torch
    dt = torch.float64
    B, nmax = len(H), max(nb)
    D = torch.zeros((ntot, ntot), dtype=dt, device=dev)
    if D0 is not None:
        for k in range(B):
            D[blk[k], blk[k]] = D0[k]
    Fp = torch.zeros((B, nmax, nmax), dtype=dt, device=dev)
    pad = torch.diag(_PAD + torch.arange(nmax, dtype=dt, device=dev))
    diis_F = [[] for _ in range(B)]
    diis_e = [[] for _ in range(B)]
    E = np.full(B, np.inf)
    done = np.zeros(B, dtype=bool) if active is None else ~np.asarray(active, dtype=bool)
    iters = np.zeros(B, dtype=int)
    C_all, eps_all = [None] * B, [None] * B

    def diis(k, F, err):
        diis_F[k].append(F)
        diis_e[k].append(err.reshape(-1))
        if len(diis_F[k]) > diis_space:
            diis_F[k].pop(0)
            diis_e[k].pop(0)
        m_ = len(diis_F[k])
        if m_ < 2:
            return F
        Ev = torch.stack(diis_e[k])
        Bm = torch.full((m_ + 1, m_ + 1), -1.0, dtype=dt, device=dev)
        Bm[:m_, :m_] = (Ev[:, None, :] * Ev[None, :, :]).sum(-1)
        Bm[-1, -1] = 0.0
        rhs = torch.zeros(m_ + 1, dtype=dt, device=dev)
        rhs[-1] = -1.0
        c = torch.linalg.solve(Bm, rhs)
        return torch.tensordot(c[:m_], torch.stack(diis_F[k]), dims=1)

    for it in range(maxit + 1):
        # molecules already done are masked out of the J/K pass
        J, K = eng.jk(D, precision=prec, fp32_threshold=fp32_threshold,
                      frag_skip=done.astype(np.int32) if done.any() else None)
        dE = np.zeros(B)
        drms = torch.zeros(B, dtype=dt, device=dev)
        Enew = torch.zeros(B, dtype=dt, device=dev)
        for k in range(B):
            if done[k]:
                continue
            b_, n = blk[k], nb[k]
            Dk = D[b_, b_]
            F = H[k] + 2.0 * J[b_, b_] - K[b_, b_]
            Enew[k] = torch.sum((F + H[k]) * Dk) + Enuc[k]
            if it > 0:
                err = A[k] @ (F @ Dk @ S[k] - S[k] @ Dk @ F) @ A[k]
                drms[k] = torch.sqrt(torch.mean(err ** 2))
                F = diis(k, F, err)
            Fp[k].copy_(pad)
            Fp[k, :n, :n] = A[k].T @ F @ A[k]
        # all molecules' Fock matrices in one batched LAPACK call on the host
        # (the padded device eigh is not used: see the incident record of
        # the batched driver)
        w_, v_ = np.linalg.eigh(Fp.cpu().numpy())
        eps, C2 = torch.as_tensor(w_, device=dev), torch.as_tensor(v_, device=dev)
        En, dr = Enew.cpu().numpy(), drms.cpu().numpy()
        for k in range(B):
            if done[k]:
                continue
            n, b_ = nb[k], blk[k]
            if it > 0:
                dE[k] = E[k] - En[k]
                iters[k] = it
                if abs(dE[k]) <= conv and dr[k] <= conv:
                    done[k] = True
                    E[k] = En[k]
                    continue
            E[k] = En[k]
            C = A[k] @ C2[k, :n, :n]
            C_all[k], eps_all[k] = C, eps[k, :n]
            Cocc = C[:, :ndocc[k]]
            D[b_, b_] = Cocc @ Cocc.T
        if done.all():
            break
    return E, D, C_all, eps_all, iters, done


def _graph_scf(eng, H, S, A, ndocc, Enuc, off, blk, ntot, conv, maxit, diis_space, prec, thr, purify_iters, dev):
    import torch
    from .scf_graph import SCF_D, SCF_F, GraphSCF
    B = len(H)
    scf = GraphSCF(eng, H, S, A, ndocc, Enuc, offsets=[int(o) for o in off[:B]], diis=True, diis_space=diis_space,
                   spectral_shift=False, conv=conv, maxit=maxit, purify_iters=purify_iters, precision=prec,
                   fp32_threshold=thr)
    try:
        scf.init([torch.zeros_like(h) for h in H])
        st = scf.run(use_graph=True)
        stalled = st["stalled"]
        bad = [k for k in range(B) if not abs(st["idempotency"][k]) < 1e-8]
        if bad:
            raise RuntimeError(f"density purification did not converge for molecules {bad}; raise purify_iters")
        Ds, Fs = scf.matrix(SCF_D), scf.matrix(SCF_F)
    finally:
        scf.close()
    D = torch.zeros((ntot, ntot), dtype=torch.float64, device=dev)
    C_all, eps_all = [], []
    for k in range(B):
        D[blk[k], blk[k]] = Ds[k]
        w_, v_ = np.linalg.eigh((A[k].T @ Fs[k] @ A[k]).cpu().numpy())
        C_all.append(A[k] @ torch.as_tensor(v_, device=dev))
        eps_all.append(torch.as_tensor(w_, device=dev))
    converged = (np.abs(st["dE"]) <= conv) & (st["drms"] <= conv)
    return list(st["E"]), D, C_all, eps_all, np.asarray(st["iterations"]), converged, stalled


def _finish(eng, one, molecules, D, C_all, eps_all, blk, aof, ndocc, E, iters, done, nb, gradient, prec,
            fp32_threshold, timing=None):
    import time
    import torch
    t0 = time.perf_counter()
    grads = None
    if gradient:
        g2 = eng.eri_grad(D, precision=prec, fp32_threshold=fp32_threshold).reshape(-1, 3)
        if timing is not None:
            torch.cuda.synchronize()
            timing["eri_grad_s"] = time.perf_counter() - t0
        grads = []
        for k, (m, e1) in enumerate(zip(molecules, one)):
            b_, nd = blk[k], ndocc[k]
            C, ek = C_all[k], eps_all[k]
            W = C[:, :nd] @ (ek[:nd, None] * C[:, :nd].T)
            g1 = e1.oei_grad(2.0 * D[b_, b_].contiguous(), (-2.0 * W).contiguous()).reshape(-1, 3)
            Z = [m.charge(i) for i in range(m.natom())]
            gk = (g1 + g2[aof[k]:aof[k + 1]]).cpu().numpy() + np.asarray(
                nuclear_repulsion_grad(m.geometry(), Z)).reshape(-1, 3)
            grads.append(gk.reshape(-1))
    for e1 in one:
        e1.close()
    eng.close()
    aux = {"iterations": iters.tolist(), "converged": done.tolist(), "nbf": nb}
    if timing is not None:
        torch.cuda.synchronize()
        aux.update(timing)
        aux["finish_s"] = time.perf_counter() - t0
    return list(E), grads, aux
