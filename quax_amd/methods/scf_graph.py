"""RHF SCF iterations replayed from one captured HIP graph (qx_scf_* in
include/quax_amd.h), for one molecule or a batch of molecules sharing one
fragmented integral context.

Same iteration as restricted_hartree_fock (quax/methods/hartree_fock.py:
66-99, with optional DIIS): J/K of the density, F = H + 2J - K, the energy
sum (F + H) D + E_nuc of the density that built F, the DIIS error
A (F D S - S D F) A, and the density of the orthogonalised Fock matrix.
That density comes from trace-correcting purification on the device instead
of an eigensolver, so one iteration is a fixed kernel sequence with no host
synchronisation: it is captured once and replayed; the host reads the
convergence flags every `check_every` replays.  Orbitals and orbital
energies are computed once, from the last Fock matrix, after convergence.
The purified density carries rounding errors that the orthogonaliser
amplifies (by up to 1/lambda_min(S)); when that floor lies above the
requested threshold the DIIS error stalls, and the last iterations run
eagerly with eigh-based densities from the graph's density.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib
from .._lib import QX_PREC_FP64, QX_PREC_MIXED

SCF_D, SCF_D_OLD, SCF_F, SCF_F_UNEXTRAPOLATED = range(4)


def _pad(mats, nm, dev):
    import torch
    out = torch.zeros((len(mats), nm, nm), dtype=torch.float64, device=dev)
    for b, m in enumerate(mats):
        n = m.shape[0]
        out[b, :n, :n] = m
    return out


class GraphSCF:
    """Native SCF state for nmol molecules of one context.  H, S, A: lists of
    nbf_b x nbf_b device tensors; offsets: first AO of each molecule in the
    context (0 for a single molecule)."""

    def __init__(self, engine, H, S, A, nocc, enuc, offsets=None, diis=True, diis_space=20,
                 spectral_shift=False, conv=1e-10, maxit=100, purify_iters=100, precision=QX_PREC_FP64,
                 fp32_threshold=0.0):
        import torch
        self.L = engine.L
        self.engine = engine
        self.dev = engine.device
        self.nbf = [int(h.shape[0]) for h in H]
        self.nmol = len(H)
        self.nm = max(self.nbf)
        self.stream = torch.cuda.Stream(device=self.dev)
        offsets = [0] * self.nmol if offsets is None else list(offsets)
        self._H, self._S, self._A = (_pad(x, self.nm, self.dev) for x in (H, S, A))
        nb = (ctypes.c_int * self.nmol)(*self.nbf)
        off = (ctypes.c_int * self.nmol)(*[int(o) for o in offsets])
        no = (ctypes.c_int * self.nmol)(*[int(o) for o in nocc])
        en = (ctypes.c_double * self.nmol)(*[float(e) for e in enuc])
        h = ctypes.c_void_p()
        self.ptr = None
        rc = self.L.qx_scf_create(engine.ctx, self.nmol, self.nm, nb, off, no, en, _lib.ptr(self._H),
                                  _lib.ptr(self._S), _lib.ptr(self._A), 1 if diis else 0, int(diis_space),
                                  1 if spectral_shift else 0, float(conv), int(maxit), int(purify_iters),
                                  int(precision), float(fp32_threshold), ctypes.byref(h))
        if rc != 0:
            raise _lib.QuaxNativeError(rc, "qx_scf_create: " + (self.L.qx_scf_last_error(None) or b"").decode())
        self.ptr = h

    def _check(self, rc, what):
        if rc != 0:
            raise _lib.QuaxNativeError(rc, f"{what}: " + (self.L.qx_scf_last_error(self.ptr) or b"").decode())

    def _sid(self):
        return ctypes.c_void_p(self.stream.cuda_stream)

    def init(self, D0):
        import torch
        Dp = _pad(D0, self.nm, self.dev)
        self.stream.wait_stream(torch.cuda.current_stream(self.dev))
        with torch.cuda.stream(self.stream):
            self._check(self.L.qx_scf_init(self.ptr, _lib.ptr(Dp), self._sid()), "qx_scf_init")
        self.stream.synchronize()

    def step(self, n, use_graph=True):
        self._check(self.L.qx_scf_step(self.ptr, int(n), 1 if use_graph else 0, self._sid()), "qx_scf_step")

    def status(self):
        B = self.nmol
        E, dE, dr, idem = (np.zeros(B) for _ in range(4))
        it, done = np.zeros(B, dtype=np.int32), np.zeros(B, dtype=np.int32)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        self._check(self.L.qx_scf_status(self.ptr, E.ctypes.data_as(dp), dE.ctypes.data_as(dp),
                                         dr.ctypes.data_as(dp), it.ctypes.data_as(ip), done.ctypes.data_as(ip),
                                         idem.ctypes.data_as(dp), self._sid()), "qx_scf_status")
        return {"E": E, "dE": dE, "drms": dr, "iterations": it, "done": done.astype(bool), "idempotency": idem}

    def matrix(self, which):
        """List of nbf_b x nbf_b tensors (one per molecule)."""
        import torch
        out = torch.empty((self.nmol, self.nm, self.nm), dtype=torch.float64, device=self.dev)
        self._check(self.L.qx_scf_matrix(self.ptr, int(which), _lib.ptr(out), self._sid()), "qx_scf_matrix")
        self.stream.synchronize()
        return [out[b, :n, :n].clone() for b, n in enumerate(self.nbf)]

    def run(self, use_graph=True, check_every=4, max_replays=None, stall_checks=3, stall_below=1e-6):
        """Iterates until every molecule is done or has stalled; returns
        status() with a "stalled" mask.  A molecule stalls when its DIIS error
        is below stall_below and its best value has not halved over the last
        stall_checks status reads: the purified density's rounding floor
        (occupied-virtual errors of ~1e-14 in the orthogonal basis, which
        D = A X A^T amplifies by up to |A|^2) is then above the threshold."""
        st = self.status()
        B = self.nmol
        hist = [[] for _ in range(B)]
        stalled = np.zeros(B, dtype=bool)
        total = 0
        while not (st["done"] | stalled).all():
            k = check_every if max_replays is None else min(check_every, max_replays - total)
            if k <= 0:
                break
            self.step(k, use_graph)
            total += k
            st = self.status()
            for b in range(B):
                if st["done"][b]:
                    continue
                hist[b].append(st["drms"][b])
                h = hist[b]
                if len(h) > stall_checks and h[-1] < stall_below and min(h[-stall_checks:]) > 0.5 * min(h[:-stall_checks]):
                    stalled[b] = True
        st["stalled"] = stalled & ~st["done"]
        return st

    def close(self):
        if self.ptr is not None:
            self.stream.synchronize()
            self.L.qx_scf_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rhf_graph(engine, nelectrons, opts, H, S, A, Enuc, D0, use_graph=True):
    """Single-molecule RHF through GraphSCF; returns (E, aux) like
    restricted_hartree_fock.  opts: its checked options."""
    import torch
    from .hartree_fock import _eigh
    ndocc = nelectrons // 2
    conv = opts["conv"] if opts["conv"] is not None else (1e-9 if opts["spectral_shift"] else 1e-10)
    prec = QX_PREC_MIXED if opts["precision"] == "mixed" else QX_PREC_FP64
    scf = GraphSCF(engine, [H], [S], [A], [ndocc], [Enuc], diis=bool(opts["diis"]), diis_space=int(opts["diis_space"]),
                   spectral_shift=bool(opts["spectral_shift"]), conv=conv, maxit=int(opts["maxit"]),
                   purify_iters=int(opts["purify_iters"]), precision=prec,
                   fp32_threshold=float(opts["fp32_threshold"]))
    try:
        scf.init([D0])
        st = scf.run(use_graph=use_graph)
        if not abs(st["idempotency"][0]) < 1e-8:
            raise RuntimeError(f"density purification did not converge (tr(X - X^2) = {st['idempotency'][0]:.2e}); "
                               f"raise the 'purify_iters' option")
        D, D_old, F = (scf.matrix(w)[0] for w in (SCF_D, SCF_D_OLD, SCF_F))
    finally:
        scf.close()
    if st["stalled"][0]:
        # the last iterations with eigh-based densities, from the graph's
        from .hartree_fock import restricted_hartree_fock
        o = dict(opts, graph=False, maxit=max(int(opts["maxit"]) - int(st["iterations"][0]), 1))
        E, aux = restricted_hartree_fock(engine, nelectrons, o, return_aux_data=True, D0=D)
        aux["iterations"] += int(st["iterations"][0])
        aux["graph_iterations"] = int(st["iterations"][0])
        return E, aux
    n = H.shape[0]
    shift = torch.zeros((n, n), dtype=torch.float64, device=H.device)
    if opts["spectral_shift"]:
        shift = torch.diag(torch.linspace(0, 1, n, dtype=torch.float64, device=H.device) * 1e-9)
    eps, C2 = _eigh(A.T @ F @ A + shift, opts["eigh"])
    C = A @ C2
    E = float(st["E"][0])
    return E, dict(D=D, D_old=D_old, C=C, eps=eps, F=F, S=S, H=H, A=A, ndocc=ndocc,
                   iterations=int(st["iterations"][0]), graph_iterations=int(st["iterations"][0]), Enuc=Enuc,
                   dE=float(st["dE"][0]), drms=float(st["drms"][0]))
