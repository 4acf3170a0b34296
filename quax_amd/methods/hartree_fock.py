"""Restricted Hartree-Fock on the GPU with integral-direct J/K.

Follows quax/methods/hartree_fock.py:8-119: Cholesky orthogonalisation
(energy_utils.py:29-36), zero (or core) density guess, optional damping of
the first 10 iterations, optional 1e-9 spectral shift, energy
E = sum (F + H) D + E_nuc, convergence on |dE| and the RMS of
A (F D S - S D F) A.  The orbital gradient is evaluated with the density that
built F; the reference pairs F with the next density, which commutes with F
by construction, so there only |dE| decides convergence and the density can
stay ~1e-8 from stationarity -- visible in gradients at the 1e-9 bar.  The reference contracts a stored nbf^4 tensor
(hartree_fock.py:79-80); here J and K come from the direct J/K kernels, so
nothing n^4 is ever stored.  DIIS extrapolation (Pulay) is an optional
addition the reference lacks (it only reports the DIIS error).
"""
from __future__ import annotations

import numpy as np

from .._lib import QX_PREC_FP64, QX_PREC_MIXED

DEFAULT_OPTIONS = {
    "maxit": 100,
    "damping": False,
    "damp_factor": 0.5,
    "guess_core": False,
    "spectral_shift": True,
    "integral_algo": "libint_core",
    "ints_tolerance": 1.0e-14,
    "freeze_core": False,
    "beta": 1.0,
    "electric_field": 0,
    # additions of this implementation
    "diis": False,
    "diis_space": 20,             # cluster10 at conv 1e-10: 53 iterations with 8 vectors, 41 with 20
    "precision": "fp64",          # or "mixed"
    "fp32_threshold": 1.0e-8,
    "conv": None,                 # override the reference's 1e-9 / 1e-10
    "eigh": "auto",               # Fock diagonalisation: "device", "host" or "auto" (host up to 512)
    "pure": False,
    "device": 0,
    "graph": False,               # iterations replayed from one captured HIP graph (qx_scf_*)
    "purify_iters": 100,          # graph SCF: most purification steps per density
}


def _eigh(Fp, where="auto", nocc=None):
    """Symmetric eigensolver of the orthogonalised Fock matrix, FP64.  For
    n <= 512 LAPACK on the host is faster than the device solver (rocSOLVER
    issues ~1000 small kernels for n = 280: ~6.4 ms against ~3.7 ms plus two
    0.6 MB copies); larger matrices stay on the device.  With `nocc` the host
    path returns only the lowest nocc eigenpairs (LAPACK's MRRR driver,
    dsyevr, ~30% faster at n = 280 with 30 of them): the SCF density needs
    the occupied orbitals only."""
    import torch
    host = where == "host" or (where == "auto" and Fp.shape[-1] <= 512)
    if not host:
        return torch.linalg.eigh(Fp)
    F = Fp.detach().cpu().numpy()
    w = v = None
    if nocc is not None and 0 < nocc < F.shape[-1]:
        try:
            import scipy.linalg
            w, v = scipy.linalg.eigh(F, subset_by_index=[0, nocc - 1], driver="evr", check_finite=False)
        except ImportError:
            pass
    if w is None:
        w, v = np.linalg.eigh(F)
    return (torch.as_tensor(w, dtype=Fp.dtype).to(Fp.device), torch.as_tensor(v, dtype=Fp.dtype).to(Fp.device))


def check_options(options):
    """core.py:19-55: known keys only, values must keep the default's type."""
    opts = dict(DEFAULT_OPTIONS)
    for k, v in (options or {}).items():
        if k not in opts:
            print(f"{k} keyword option not recognized.")
            continue
        d = DEFAULT_OPTIONS[k]
        if d is None or type(v) == type(d) or (isinstance(d, float) and isinstance(v, float)):
            opts[k] = v
        else:
            print(f"Value '{v}' for keyword option '{k}' not recognized. Ignoring.")
    return opts


def nuclear_repulsion(geom, Z):
    """energy_utils.py:6-15."""
    g = np.asarray(geom).reshape(-1, 3)
    e = 0.0
    for i in range(len(Z)):
        for j in range(i):
            e += Z[i] * Z[j] / np.linalg.norm(g[i] - g[j])
    return float(e)


def nuclear_repulsion_grad(geom, Z):
    g = np.asarray(geom).reshape(-1, 3)
    out = np.zeros_like(g)
    for i in range(len(Z)):
        for j in range(len(Z)):
            if i != j:
                r = g[i] - g[j]
                out[i] -= Z[i] * Z[j] * r / np.linalg.norm(r) ** 3
    return out.reshape(-1)


def restricted_hartree_fock(engine, nelectrons, options=None, return_aux_data=False, jk=None, D0=None):
    """Returns E_scf, or (E_scf, aux) with aux holding D, C, eps, F, S, H,
    iteration count.  ``jk(D) -> (J, K)`` replaces the engine's direct J/K
    (used by tests to compare contraction paths); ``D0`` is a starting
    density (nbf x nbf) in place of the zero / core guess."""
    import torch
    opts = check_options(options)
    dev = engine.device
    dt = torch.float64
    ndocc = nelectrons // 2
    S, T, V = engine.oei()
    H = T + V
    n = S.shape[0]
    A = torch.linalg.inv(torch.linalg.cholesky(S)).T.contiguous()
    mol = engine.molecule
    Z = [mol.charge(i) for i in range(mol.natom())]
    Enuc = nuclear_repulsion(mol.geometry(), Z)
    shift = torch.zeros((n, n), dtype=dt, device=dev)
    if opts["spectral_shift"]:
        shift = torch.diag(torch.linspace(0, 1, n, dtype=dt, device=dev) * 1e-9)
    conv = opts["conv"] if opts["conv"] is not None else (1e-9 if opts["spectral_shift"] else 1e-10)
    prec = QX_PREC_MIXED if opts["precision"] == "mixed" else QX_PREC_FP64
    thr = float(opts["fp32_threshold"])

    if opts["graph"]:
        if jk is not None or opts["damping"]:
            raise ValueError("the graph SCF has no custom J/K and no damping")
        from .scf_graph import rhf_graph
        if D0 is not None:
            Dg = torch.as_tensor(D0, dtype=dt).to(dev)
            if tuple(Dg.shape) != tuple(H.shape):
                raise ValueError("D0 must be nbf x nbf")
        else:
            Dg = H.clone() if opts["guess_core"] else torch.zeros_like(H)
        E, aux = rhf_graph(engine, nelectrons, opts, H, S, A, Enuc, Dg)
        return (E, aux) if return_aux_data else E

    def build_jk(D):
        if jk is not None:
            return jk(D)
        return engine.jk(D, precision=prec, fp32_threshold=thr)

    def rhf_iter(F, D, full=False):
        E = torch.sum((F + H) * D) + Enuc
        Fp = A.T @ F @ A + shift
        eps, C2 = _eigh(Fp, opts["eigh"], nocc=None if full else ndocc)
        C = A @ C2
        Cocc = C[:, :ndocc]
        return E, Cocc @ Cocc.T, C, eps

    diis_F, diis_e = [], []

    def diis_extrapolate(F, err):
        diis_F.append(F)
        diis_e.append(err)
        if len(diis_F) > opts["diis_space"]:
            diis_F.pop(0)
            diis_e.pop(0)
        m = len(diis_F)
        if m < 2:
            return F
        E = torch.stack([e.reshape(-1) for e in diis_e])
        B = torch.full((m + 1, m + 1), -1.0, dtype=dt, device=dev)
        # one fused reduction: a GEMM with m x m output and n^2-long inner
        # dimension maps onto a single rocBLAS tile and takes milliseconds
        B[:m, :m] = (E[:, None, :] * E[None, :, :]).sum(-1)
        B[-1, -1] = 0.0
        rhs = torch.zeros(m + 1, dtype=dt, device=dev)
        rhs[-1] = -1.0
        try:
            c = torch.linalg.solve(B, rhs)
        except Exception:
            return F
        return torch.tensordot(c[:m], torch.stack(diis_F), dims=1)

    if D0 is not None:
        D = torch.as_tensor(D0, dtype=dt).to(dev).clone()
        if tuple(D.shape) != tuple(H.shape):
            raise ValueError("D0 must be nbf x nbf")
    else:
        D = H.clone() if opts["guess_core"] else torch.zeros_like(H)
    J, K = build_jk(D)
    F = H + 2 * J - K
    D_old = D.clone()
    E, D, C, eps = rhf_iter(F, D)
    it = 0
    dE, drms = 1.0, 1.0
    while (abs(dE) > conv or drms > conv) and it < opts["maxit"]:
        if opts["damping"] and it < 10:
            D = D_old * opts["damp_factor"] + D * opts["damp_factor"]
        D_old = D.clone()
        J, K = build_jk(D)
        F = H + 2 * J - K
        Fu = F
        # orbital gradient A (F D S - S D F) A of the density D that built F
        err = A @ (F @ D @ S - S @ D @ F) @ A
        if opts["diis"]:
            F = diis_extrapolate(F, err)
        E_new, D, C, eps = rhf_iter(F, D_old)
        if opts["diis"]:
            # energy of the density that built the (unextrapolated) Fock matrix
            E_new = torch.sum((Fu + H) * D_old) + Enuc
        drms = float(torch.sqrt(torch.mean(err ** 2)))
        dE = float(E - E_new)
        E = E_new
        it += 1
    E = float(E)
    if not return_aux_data:
        return E
    # every orbital and orbital energy of the last Fock matrix (the loop
    # solved for the occupied ones only)
    _, _, C, eps = rhf_iter(F, D_old, full=True)
    return E, dict(D=D, D_old=D_old, C=C, eps=eps, F=F, S=S, H=H, A=A, ndocc=ndocc, iterations=it, Enuc=Enuc,
                   dE=dE, drms=drms)


def rhf_gradient(engine, aux, precision="fp64", fp32_threshold=1e-8):
    """Analytic RHF gradient consistent with E = sum (F+H) D + E_nuc:
    dE/dR = 2 D.dH + sum (2 D_pq D_rs - D_pr D_qs) d(pq|rs) - 2 W.dS + dE_nuc,
    W = C_occ diag(eps_occ) C_occ^T.  The two-electron term is the adjoint
    kernel (integrals recomputed, no derivative tensor stored); the
    one-electron term is contracted on the fly as well."""
    D, C, eps, nd = aux["D"], aux["C"], aux["eps"], aux["ndocc"]
    W = C[:, :nd] @ (eps[:nd, None] * C[:, :nd].T)
    prec = QX_PREC_MIXED if precision == "mixed" else QX_PREC_FP64
    g2 = engine.eri_grad(D, precision=prec, fp32_threshold=fp32_threshold)
    g1 = engine.oei_grad(2.0 * D, -2.0 * W)
    mol = engine.molecule
    Z = [mol.charge(i) for i in range(mol.natom())]
    gn = nuclear_repulsion_grad(mol.geometry(), Z)
    return (g1 + g2).cpu().numpy() + gn


def nuclear_repulsion_hessian(geom, Z):
    """(3N, 3N) second derivatives of sum_{i<j} Z_i Z_j / |R_i - R_j|."""
    g = np.asarray(geom, dtype=np.float64).reshape(-1, 3)
    n = g.shape[0]
    Hn = np.zeros((3 * n, 3 * n))
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            r = g[i] - g[j]
            d = np.linalg.norm(r)
            blk = Z[i] * Z[j] * (3.0 * np.outer(r, r) / d ** 5 - np.eye(3) / d ** 3)
            Hn[3 * i:3 * i + 3, 3 * i:3 * i + 3] += blk
            Hn[3 * i:3 * i + 3, 3 * j:3 * j + 3] -= blk
    return Hn


def _cphf_solve(jk2, Cv, Co, denom, rhs, conv, maxit):
    """Preconditioned conjugate gradients for the RHF orbital response
    (eps_a - eps_i) U_ai + [Cv^T G(Cv U Co^T + Co U^T Cv^T) Co]_ai = rhs_ai,
    G(X) = 2 J(X) - K(X): the (symmetric, positive definite at a stable
    minimum) RHF orbital Hessian, one J/K build per iteration."""
    import torch

    def op(U):
        X = Cv @ U @ Co.T
        return denom * U + Cv.T @ jk2(X + X.T) @ Co

    U = rhs / denom
    r = rhs - op(U)
    z = r / denom
    p = z.clone()
    rz = torch.sum(r * z)
    bnorm = float(torch.linalg.norm(rhs)) or 1.0
    for it in range(maxit):
        if float(torch.linalg.norm(r)) <= conv * bnorm:
            return U, it
        Ap = op(p)
        alpha = rz / torch.sum(p * Ap)
        U = U + alpha * p
        r = r - alpha * Ap
        z = r / denom
        rz_new = torch.sum(r * z)
        p = z + (rz_new / rz) * p
        rz = rz_new
    raise RuntimeError(f"CPHF did not converge in {maxit} iterations "
                       f"(residual {float(torch.linalg.norm(r)) / bnorm:.2e})")


def rhf_hessian_vector_product(engine, aux, v, cphf_conv=1e-10, cphf_maxit=200):
    """H v for the converged RHF energy of ``aux`` (restricted_hartree_fock's
    aux data), by forward-over-reverse: the total derivative along the
    geometry direction v of the analytic gradient
        g = 2 D.dH + sum (2 D_pq D_rs - D_pr D_qs) d(pq|rs) - 2 W.dS + dE_nuc,
    with W = D F D.  The density response comes from the coupled-perturbed
    equations (skeleton derivative Fock matrix from the derivative integrals'
    J/K along v); the terms that differentiate integral derivatives run the
    adjoint kernels in forward mode (qx_eri_hess_dir, qx_oei_hess_dir), and
    the response of the two-electron term uses the exact polarisation
    identity of the quadratic Gamma(D):
        sum dGamma d(pq|rs) = [grad2(D + dD) - grad2(D - dD)] / 2.
    No second-derivative tensor is formed.  Returns a (3N,) numpy array."""
    import torch
    dev = engine.device
    dt = torch.float64
    natom = engine.natom
    v = np.asarray(v, dtype=np.float64).reshape(-1)
    if v.shape != (3 * natom,):
        raise ValueError("v must have 3 * natom entries")
    vt = torch.as_tensor(v, dtype=dt, device=dev)
    D, C, eps, nd, S, H = aux["D"], aux["C"], aux["eps"], aux["ndocc"], aux["S"], aux["H"]

    def jk2(X):
        J, K = engine.jk(X)
        return 2.0 * J - K

    F = H + jk2(D)
    W = D @ F @ D
    dS, dT, dV = engine.oei_deriv()
    Sv = torch.tensordot(vt, dS, dims=1)
    Hv = torch.tensordot(vt, dT + dV, dims=1)
    Jv, Kv = engine.jk_deriv_dir(D, vt)
    Gv = 2.0 * Jv - Kv
    Co, Cv = C[:, :nd], C[:, nd:]
    eo, ev = eps[:nd], eps[nd:]
    St = C.T @ Sv @ C
    D0 = -Co @ St[:nd, :nd] @ Co.T
    F0 = C.T @ (Hv + Gv + jk2(D0)) @ C
    rhs = -(F0[nd:, :nd] - St[nd:, :nd] * eo[None, :])
    denom = ev[:, None] - eo[None, :]
    U, _ = _cphf_solve(jk2, Cv, Co, denom, rhs, cphf_conv, cphf_maxit)
    X = Cv @ U @ Co.T
    Dd = D0 + X + X.T
    Fd = Hv + Gv + jk2(Dd)
    Wd = Dd @ F @ D + D @ Fd @ D + D @ F @ Dd
    h = engine.oei_grad(2.0 * Dd, -2.0 * Wd)
    h = h + engine.oei_hess_dir(2.0 * D, -2.0 * W, vt)
    h = h + 0.5 * (engine.eri_grad(D + Dd) - engine.eri_grad(D - Dd))
    h = h + engine.eri_hess_dir(D, vt)
    mol = engine.molecule
    Z = [mol.charge(i) for i in range(mol.natom())]
    return h.cpu().numpy() + nuclear_repulsion_hessian(mol.geometry(), Z) @ v
