"""AO -> MO two-electron integral transformations on the GPU.

Mirrors quax/methods/energy_utils.py:46-73: ``tei_transformation(G, C)`` and
``partial_tei_transformation(G, Ci, Cj, Ck, Cl)`` return

    out[I, J, K, L] = sum_pqrs Ci[p, I] Cj[q, J] Ck[r, K] Cl[s, L] G[p, q, r, s]

computed as the reference does, four contractions of the last index (l, then
k, j, i), each a hand-written FP64 MFMA GEMM (qx_partial_tei_transformation in
libquax_hip.so).  Inputs are torch tensors or arrays; G and the coefficients
are moved to the GPU as FP64 and the result stays there.  There is no CPU path.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib


def _dev(torch, x, device):
    return torch.as_tensor(x, dtype=torch.float64).to(device).contiguous()


def partial_tei_transformation(G, Ci, Cj, Ck, Cl, device=None):
    import torch
    if not torch.cuda.is_available():
        raise _lib.NativeUnavailable("quax_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    L = _lib.load()
    if device is None:
        device = G.device if isinstance(G, torch.Tensor) and G.is_cuda else torch.device("cuda", 0)
    G = _dev(torch, G, device)
    Cs = [_dev(torch, C, device) for C in (Ci, Cj, Ck, Cl)]
    if G.dim() != 4:
        raise ValueError("G must be a 4-index tensor")
    n = np.array(G.shape, dtype=np.int32)
    for k, C in enumerate(Cs):
        if C.dim() != 2 or C.shape[0] != n[k]:
            raise ValueError(f"coefficient matrix {k} must have {n[k]} rows")
    m = np.array([C.shape[1] for C in Cs], dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    nw = L.qx_tei_transform_work(n.ctypes.data_as(ip), m.ctypes.data_as(ip))
    if nw < 0:
        raise ValueError("tei transformation: every dimension must be positive")
    work = torch.empty(int(nw), dtype=torch.float64, device=device)
    out = torch.empty(tuple(int(x) for x in m), dtype=torch.float64, device=device)
    with torch.cuda.device(device):
        stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
        rc = L.qx_partial_tei_transformation(_lib.ptr(G), n.ctypes.data_as(ip), *(_lib.ptr(C) for C in Cs),
                                             m.ctypes.data_as(ip), _lib.ptr(out), _lib.ptr(work), stream)
    if rc != _lib.QX_OK:
        raise _lib.QuaxNativeError(rc, L.qx_gemm_last_error().decode())
    return out


def tei_transformation(G, C, device=None):
    """All four indices with the same coefficients (energy_utils.py:46-55)."""
    return partial_tei_transformation(G, C, C, C, C, device=device)
