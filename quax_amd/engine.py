"""Device-resident ERI engine: one molecule + basis on one GPU.

Wraps a ``qx_ctx`` of libquax_hip.so.  Buffers are torch tensors on the
engine's device; every call goes through the C ABI and runs hand-written HIP
kernels.  There is no CPU fallback: without the library or a GPU the
constructor raises.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import QX_PREC_FP64, QX_PREC_MIXED, NativeUnavailable, check, ptr


def _require_gpu():
    import torch
    if not torch.cuda.is_available():
        raise NativeUnavailable("quax_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


class EriEngine:
    def __init__(self, molecule, basis, ints_tol: float = 1e-14, device: int = 0, fragments=None):
        """fragments: optional fragment id per atom (qx_initialize_fragments):
        no shell pair or quartet spans two fragments."""
        torch = _require_gpu()
        self.torch = torch
        self.L = _lib.load()
        self.molecule = molecule
        self.basis = basis
        self.device = torch.device("cuda", device)
        self.natom = molecule.natom()
        xyz = np.ascontiguousarray(molecule.geometry().reshape(-1), dtype=np.float64)
        Z = np.array([molecule.charge(i) for i in range(self.natom)], dtype=np.float64)
        l, atom, nprim, exps, coefs = basis.arrays()
        self._keep = (xyz, Z, l, atom, nprim, exps, coefs)
        ctx = ctypes.c_void_p()
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        with torch.cuda.device(self.device):
            torch.cuda.current_stream().synchronize()
            args = (self.natom, xyz.ctypes.data_as(dp), Z.ctypes.data_as(dp), len(l), l.ctypes.data_as(ip),
                    atom.ctypes.data_as(ip), nprim.ctypes.data_as(ip), exps.ctypes.data_as(dp),
                    coefs.ctypes.data_as(dp), 1 if basis.pure else 0, float(ints_tol), device)
            if fragments is None:
                rc = self.L.qx_initialize(*args, ctypes.byref(ctx))
            else:
                fr = np.ascontiguousarray(fragments, dtype=np.int32)
                if fr.shape != (self.natom,):
                    raise ValueError("fragments needs one id per atom")
                self._keep = self._keep + (fr,)
                rc = self.L.qx_initialize_fragments(*args, fr.ctypes.data_as(ip), ctypes.byref(ctx))
        if rc != 0:
            raise _lib.QuaxNativeError(rc, "qx_initialize: " + self.L.qx_last_error(None).decode())
        self.ctx = ctx
        nbf, nc = ctypes.c_int(), ctypes.c_int()
        check(self.L.qx_nbf(self.ctx, ctypes.byref(nbf), ctypes.byref(nc)), self.ctx)
        self.nbf, self.ncart = nbf.value, nc.value
        self.ints_tol = ints_tol

    # ------------------------------------------------------------ lifetime
    def close(self):
        if getattr(self, "ctx", None):
            self.L.qx_finalize(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ helpers
    def _empty(self, *shape):
        return self.torch.empty(shape, dtype=self.torch.float64, device=self.device)

    def _zeros(self, *shape):
        return self.torch.zeros(shape, dtype=self.torch.float64, device=self.device)

    def _stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def _dev(self, x):
        t = self.torch.as_tensor(x, dtype=self.torch.float64, device=self.device)
        return t.contiguous()

    # ------------------------------------------------------------ queries
    def schwarz_screening(self):
        ns = self.basis.nshell()
        vals = np.zeros((ns, ns))
        pairs = np.zeros(ns * (ns + 1), dtype=np.int32)
        npairs = ctypes.c_int()
        check(self.L.qx_schwarz_screening(self.ctx, vals.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                          pairs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                          ctypes.byref(npairs)), self.ctx)
        return vals, pairs[:2 * npairs.value].reshape(-1, 2)

    def num_quartets(self) -> int:
        n = ctypes.c_longlong()
        check(self.L.qx_num_quartets(self.ctx, ctypes.byref(n)), self.ctx)
        return n.value

    def class_quartets(self):
        """Per kernel class: ((la,lb,lc,ld), quartets, non-single-atom quartets, derivative parts)."""
        nc = self.L.qx_num_classes()
        eri = (ctypes.c_longlong * nc)()
        deri = (ctypes.c_longlong * nc)()
        check(self.L.qx_class_quartets(self.ctx, eri, deri), self.ctx)
        out = []
        for k in range(nc):
            l4 = (ctypes.c_int * 4)()
            npart = ctypes.c_int()
            self.L.qx_class_info(k, l4, ctypes.byref(npart))
            out.append((tuple(l4), eri[k], deri[k], npart.value))
        return out

    def mixed_quartets(self, fp32_threshold: float, pass_=-1):
        """(FP64, FP32) contracted quartet counts of a mixed-precision pass at
        the threshold: pass_ 0 = J/K, 1 = gradient (the classes whose FP32
        policy admits FP32), -1 = every class below the threshold."""
        n64, n32 = ctypes.c_longlong(), ctypes.c_longlong()
        check(self.L.qx_mixed_pass_quartets(self.ctx, float(fp32_threshold), int(pass_), ctypes.byref(n64),
                                            ctypes.byref(n32)), self.ctx, "mixed_quartets")
        return n64.value, n32.value

    def fp32_classes(self, pass_: int):
        """Per kernel class, whether the mixed pass (0 J/K, 1 gradient) may use FP32."""
        nc = self.L.qx_num_classes()
        f = np.zeros(nc, dtype=np.int32)
        check(self.L.qx_get_fp32_classes(self.ctx, int(pass_), f.ctypes.data_as(ctypes.c_void_p)), self.ctx)
        return f.astype(bool)

    def class_tiles(self, pass_: int):
        """Per kernel class, threads per block of the streaming kernels of a
        pass (0 J/K, 1 adjoint gradient)."""
        f = np.zeros(self.L.qx_num_classes(), dtype=np.int32)
        check(self.L.qx_get_class_tiles(self.ctx, int(pass_), f.ctypes.data_as(ctypes.c_void_p)), self.ctx)
        return f

    def set_class_tiles(self, pass_: int, block=None):
        """Sets the per-class threads per block (multiples of 64 up to 256;
        None: the built-in tuned table)."""
        if block is None:
            check(self.L.qx_set_class_tiles(self.ctx, int(pass_), None), self.ctx)
            return
        f = np.ascontiguousarray(np.asarray(block, dtype=np.int32))
        if f.shape != (self.L.qx_num_classes(),):
            raise ValueError("one block size per kernel class")
        check(self.L.qx_set_class_tiles(self.ctx, int(pass_), f.ctypes.data_as(ctypes.c_void_p)), self.ctx,
              "set_class_tiles")

    def grad_wave_reduce(self):
        """Per kernel class, whether the adjoint-gradient kernels reduce
        wave-uniform centres across the wavefront."""
        f = np.zeros(self.L.qx_num_classes(), dtype=np.int32)
        check(self.L.qx_get_grad_wave_reduce(self.ctx, f.ctypes.data_as(ctypes.c_void_p)), self.ctx)
        return f.astype(bool)

    def set_grad_wave_reduce(self, flags=None):
        """Sets it per class (None: the built-in tuned table)."""
        if flags is None:
            check(self.L.qx_set_grad_wave_reduce(self.ctx, None), self.ctx)
            return
        f = np.ascontiguousarray(np.asarray(flags, dtype=np.int32))
        if f.shape != (self.L.qx_num_classes(),):
            raise ValueError("one flag per kernel class")
        check(self.L.qx_set_grad_wave_reduce(self.ctx, f.ctypes.data_as(ctypes.c_void_p)), self.ctx)

    def grad_copies(self):
        """Per kernel class, copies of the gradient accumulators its adjoint
        blocks flush into."""
        f = np.zeros(self.L.qx_num_classes(), dtype=np.int32)
        check(self.L.qx_get_grad_copies(self.ctx, f.ctypes.data_as(ctypes.c_void_p)), self.ctx)
        return f

    def set_grad_copies(self, copies=None):
        """Sets them per class (1..64; None: the built-in tuned table)."""
        if copies is None:
            check(self.L.qx_set_grad_copies(self.ctx, None), self.ctx)
            return
        f = np.ascontiguousarray(np.asarray(copies, dtype=np.int32))
        if f.shape != (self.L.qx_num_classes(),):
            raise ValueError("one copy count per kernel class")
        check(self.L.qx_set_grad_copies(self.ctx, f.ctypes.data_as(ctypes.c_void_p)), self.ctx, "set_grad_copies")

    def jk_rows(self):
        """Per kernel class, whether its J/K kernels stage K' rows in LDS."""
        f = np.zeros(self.L.qx_num_classes(), dtype=np.int32)
        check(self.L.qx_get_jk_rows(self.ctx, f.ctypes.data_as(ctypes.c_void_p)), self.ctx)
        return f

    def set_jk_rows(self, flags=None):
        """Sets them per class (None: all on)."""
        if flags is None:
            check(self.L.qx_set_jk_rows(self.ctx, None), self.ctx)
            return
        f = np.ascontiguousarray(np.asarray(flags, dtype=np.int32))
        if f.shape != (self.L.qx_num_classes(),):
            raise ValueError("one flag per kernel class")
        check(self.L.qx_set_jk_rows(self.ctx, f.ctypes.data_as(ctypes.c_void_p)), self.ctx, "set_jk_rows")

    def grad_window(self):
        """Widest atom window of the adjoint kernels' LDS gradient accumulators."""
        a = ctypes.c_int(0)
        check(self.L.qx_get_grad_window(self.ctx, ctypes.byref(a)), self.ctx)
        return a.value

    def set_grad_window(self, atoms: int = 0):
        """Sets it (0..256 atoms; 0: the default, 256)."""
        check(self.L.qx_set_grad_window(self.ctx, int(atoms)), self.ctx, "set_grad_window")

    def set_fp32_classes(self, pass_: int, flags=None):
        """Sets the per-class FP32 policy of a mixed pass (None: built-in default)."""
        if flags is None:
            check(self.L.qx_set_fp32_classes(self.ctx, int(pass_), None), self.ctx)
            return
        f = np.ascontiguousarray(np.asarray(flags, dtype=np.int32))
        if f.shape != (self.L.qx_num_classes(),):
            raise ValueError("one flag per kernel class")
        check(self.L.qx_set_fp32_classes(self.ctx, int(pass_), f.ctypes.data_as(ctypes.c_void_p)), self.ctx)

    def set_streams(self, n: int):
        check(self.L.qx_set_streams(self.ctx, int(n)), self.ctx, "set_streams")

    def profile(self, on: bool):
        check(self.L.qx_profile_enable(self.ctx, 1 if on else 0), self.ctx)

    def profile_read(self):
        """(ms, launches, primitive quartets) per profile slot."""
        ns = self.L.qx_profile_slots()
        ms = (ctypes.c_double * ns)()
        la = (ctypes.c_longlong * ns)()
        pq = (ctypes.c_longlong * ns)()
        check(self.L.qx_profile_read(self.ctx, ms, la, pq), self.ctx)
        return np.array(ms[:]), np.array(la[:]), np.array(pq[:])

    # ------------------------------------------------------------ tensors
    def compute_2e_int(self):
        n = self.nbf
        out = self._empty(n, n, n, n)
        check(self.L.qx_compute_2e_int(self.ctx, ptr(out), self._stream()), self.ctx, "compute_2e_int")
        return out

    def eri_deriv_core(self, deriv_order: int = 1):
        n = self.nbf
        nu = self.L.qx_how_many_derivs(self.natom, deriv_order)
        out = self._empty(nu, n, n, n, n)
        check(self.L.qx_eri_deriv_core(self.ctx, int(deriv_order), ptr(out), self._stream()), self.ctx,
              "eri_deriv_core")
        return out

    def compute_2e_deriv(self, deriv_vec):
        dv = np.ascontiguousarray(np.asarray(deriv_vec).reshape(-1), dtype=np.int32)
        n = self.nbf
        out = self._empty(n, n, n, n)
        check(self.L.qx_compute_2e_deriv(self.ctx, dv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(dv),
                                         ptr(out), self._stream()), self.ctx, "compute_2e_deriv")
        return out

    def store(self):
        nd = ctypes.c_longlong()
        check(self.L.qx_store_size(self.ctx, ctypes.byref(nd)), self.ctx)
        out = self._empty(nd.value)
        check(self.L.qx_eri_store(self.ctx, ptr(out), self._stream()), self.ctx, "eri_store")
        return out

    def store_map(self):
        """Layout of ``store()``: a list of dicts, one per quartet list, with
        the class ``l4``, element offset ``base``, quartet count ``nquart``
        and, per quartet in stored order, the bra and ket shell pairs
        (``p1``, ``p2``: indices into ``pairs``); plus ``pairs``, the
        oriented (sa, sb) shells of every screened pair."""
        nc = self.L.qx_num_classes()
        nl = ctypes.c_int()
        l4 = (ctypes.c_int * (4 * nc))()
        base = (ctypes.c_longlong * nc)()
        nq = (ctypes.c_longlong * nc)()
        check(self.L.qx_store_lists(self.ctx, ctypes.byref(nl), l4, base, nq), self.ctx, "store_lists")
        npair = ctypes.c_int()
        check(self.L.qx_pair_shells(self.ctx, ctypes.byref(npair), None, None), self.ctx, "pair_shells")
        sa = np.zeros(npair.value, dtype=np.int32)
        sb = np.zeros(npair.value, dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        check(self.L.qx_pair_shells(self.ctx, ctypes.byref(npair), sa.ctypes.data_as(ip), sb.ctypes.data_as(ip)),
              self.ctx, "pair_shells")
        lists = []
        for k in range(nl.value):
            p1 = np.zeros(nq[k], dtype=np.int32)
            p2 = np.zeros(nq[k], dtype=np.int32)
            check(self.L.qx_store_quartet_pairs(self.ctx, k, p1.ctypes.data_as(ip), p2.ctypes.data_as(ip)),
                  self.ctx, "store_quartet_pairs")
            lists.append(dict(l4=tuple(l4[4 * k:4 * k + 4]), base=base[k], nquart=nq[k], p1=p1, p2=p2))
        return lists, np.stack([sa, sb], axis=1)

    # ------------------------------------------------------------ direct
    def jk(self, D, precision: int = QX_PREC_FP64, fp32_threshold: float = 0.0, out=None, frag_skip=None):
        """J and K of D.  frag_skip (batch contexts): per fragment id, nonzero
        = leave that fragment's quartets out (its blocks come out zero)."""
        D = self._dev(D)
        n = self.nbf
        J, K = (out if out is not None else (self._empty(n, n), self._empty(n, n)))
        if frag_skip is None:
            check(self.L.qx_jk_build(self.ctx, ptr(D), ptr(J), ptr(K), int(precision), float(fp32_threshold),
                                     self._stream()), self.ctx, "jk_build")
        else:
            sk = self.torch.as_tensor(np.asarray(frag_skip, dtype=np.int32), device=self.device)
            check(self.L.qx_jk_build_masked(self.ctx, ptr(D), ptr(J), ptr(K), int(precision),
                                            float(fp32_threshold), ptr(sk), self._stream()), self.ctx,
                  "jk_build_masked")
        return J, K

    def eri_grad(self, D, precision: int = QX_PREC_FP64, fp32_threshold: float = 0.0):
        D = self._dev(D)
        g = self._zeros(3 * self.natom)
        check(self.L.qx_eri_grad(self.ctx, ptr(D), ptr(g), int(precision), float(fp32_threshold), self._stream()),
              self.ctx, "eri_grad")
        return g

    # ------------------------------------------------------------ forward over reverse
    def eri_hess_dir(self, D, v):
        """sum_pqrs Gamma(D)_pqrs d/dv d(pq|rs)/dR (3N): the adjoint gradient
        kernels in forward mode along the geometry direction v."""
        D, v = self._dev(D), self._dev(v)
        g = self._zeros(3 * self.natom)
        check(self.L.qx_eri_hess_dir(self.ctx, ptr(D), ptr(v), ptr(g), self._stream()), self.ctx, "eri_hess_dir")
        return g

    def jk_deriv_dir(self, D, v):
        """J and K of the integrals' directional derivative along v at fixed D."""
        D, v = self._dev(D), self._dev(v)
        n = self.nbf
        J, K = self._empty(n, n), self._empty(n, n)
        check(self.L.qx_jk_deriv_dir(self.ctx, ptr(D), ptr(v), ptr(J), ptr(K), self._stream()), self.ctx,
              "jk_deriv_dir")
        return J, K

    def oei_hess_dir(self, Pd, Wd, v):
        """sum_pq Pd d/dv d(T+V)/dR + Wd d/dv dS/dR (3N)."""
        Pd, Wd, v = self._dev(Pd), self._dev(Wd), self._dev(v)
        g = self._zeros(3 * self.natom)
        check(self.L.qx_oei_hess_dir(self.ctx, ptr(Pd), ptr(Wd), ptr(v), ptr(g), self._stream()), self.ctx,
              "oei_hess_dir")
        return g

    def oei(self):
        n = self.nbf
        S, T, V = self._empty(n, n), self._empty(n, n), self._empty(n, n)
        check(self.L.qx_oei(self.ctx, ptr(S), ptr(T), ptr(V), self._stream()), self.ctx, "oei")
        return S, T, V

    def oei_grad(self, Pd, Wd):
        Pd, Wd = self._dev(Pd), self._dev(Wd)
        g = self._zeros(3 * self.natom)
        check(self.L.qx_oei_grad(self.ctx, ptr(Pd), ptr(Wd), ptr(g), self._stream()), self.ctx, "oei_grad")
        return g

    def oei_deriv(self):
        n = self.nbf
        shp = (3 * self.natom, n, n)
        dS, dT, dV = self._empty(*shp), self._empty(*shp), self._empty(*shp)
        check(self.L.qx_oei_deriv(self.ctx, ptr(dS), ptr(dT), ptr(dV), self._stream()), self.ctx, "oei_deriv")
        return dS, dT, dV
