"""ctypes binding of libquax_hip.so (C ABI in include/quax_amd.h).

The product path has no CPU fallback: if the native library is missing or
no GPU is visible, every entry point raises ``NativeUnavailable``.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# QUAX_AMD_LIB selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("QUAX_AMD_LIB") or os.path.join(_HERE, "lib", "libquax_hip.so")

QX_OK, QX_EINVAL, QX_ENOMEM, QX_EHIP, QX_EUNSUPPORTED, QX_ENODEVICE = range(6)
QX_PREC_FP64, QX_PREC_MIXED = 0, 1

_lib = None


class NativeUnavailable(RuntimeError):
    """The HIP library is not built or cannot run here."""


class QuaxNativeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[quax_amd code {code}] {msg}")
        self.code = code


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_ll = ctypes.c_longlong

# name -> (restype, argtypes)
SIGNATURES = {
    "qx_initialize": (_i, [_i, _dp, _dp, _i, _ip, _ip, _ip, _dp, _dp, _i, ctypes.c_double, _i,
                           ctypes.POINTER(_vp)]),
    "qx_initialize_fragments": (_i, [_i, _dp, _dp, _i, _ip, _ip, _ip, _dp, _dp, _i, ctypes.c_double, _i, _ip,
                                     ctypes.POINTER(_vp)]),
    "qx_finalize": (_i, [_vp]),
    "qx_last_error": (ctypes.c_char_p, [_vp]),
    "qx_nbf": (_i, [_vp, _ip, _ip]),
    "qx_schwarz_screening": (_i, [_vp, _dp, _ip, _ip]),
    "qx_num_quartets": (_i, [_vp, ctypes.POINTER(_ll)]),
    "qx_compute_2e_int": (_i, [_vp, _vp, _vp]),
    "qx_eri_deriv_core": (_i, [_vp, _i, _vp, _vp]),
    "qx_compute_2e_deriv": (_i, [_vp, _ip, _i, _vp, _vp]),
    "qx_how_many_derivs": (_ll, [_i, _i]),
    "qx_generate_multi_index_lookup": (_ll, [_i, _i, _ip, _ll]),
    "qx_jk_build": (_i, [_vp, _vp, _vp, _vp, _i, ctypes.c_double, _vp]),
    "qx_jk_build_noalloc": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "qx_jk_noalloc_capacity": (_ll, []),
    "qx_eri_grad": (_i, [_vp, _vp, _vp, _i, ctypes.c_double, _vp]),
    "qx_eri_hess_dir": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "qx_jk_deriv_dir": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "qx_oei_hess_dir": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "qx_oei": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "qx_oei_grad": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "qx_oei_deriv": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "qx_store_size": (_i, [_vp, ctypes.POINTER(_ll)]),
    "qx_eri_store": (_i, [_vp, _vp, _vp]),
    "qx_store_lists": (_i, [_vp, _ip, _ip, ctypes.POINTER(_ll), ctypes.POINTER(_ll)]),
    "qx_store_quartet_pairs": (_i, [_vp, _i, _ip, _ip]),
    "qx_pair_shells": (_i, [_vp, _ip, _ip, _ip]),
    "qx_tei_transform_work": (_ll, [_ip, _ip]),
    "qx_partial_tei_transformation": (_i, [_vp, _ip, _vp, _vp, _vp, _vp, _ip, _vp, _vp, _vp]),
    "qx_gemm_last_error": (ctypes.c_char_p, []),
    "qx_num_classes": (_i, []),
    "qx_profile_slots": (_i, []),
    "qx_profile_slot": (_i, [_i, _ip, _ip, _ip, _ip]),
    "qx_mixed_quartets": (_i, [_vp, ctypes.c_double, ctypes.POINTER(_ll), ctypes.POINTER(_ll)]),
    "qx_mixed_pass_quartets": (_i, [_vp, ctypes.c_double, _i, ctypes.POINTER(_ll), ctypes.POINTER(_ll)]),
    "qx_set_fp32_classes": (_i, [_vp, _i, _vp]),
    "qx_get_fp32_classes": (_i, [_vp, _i, _vp]),
    "qx_set_streams": (_i, [_vp, _i]),
    "qx_set_class_tiles": (_i, [_vp, _i, _vp]),
    "qx_set_grad_wave_reduce": (_i, [_vp, _vp]),
    "qx_set_grad_copies": (_i, [_vp, _vp]),
    "qx_get_grad_copies": (_i, [_vp, _vp]),
    "qx_set_jk_rows": (_i, [_vp, _vp]),
    "qx_get_jk_rows": (_i, [_vp, _vp]),
    "qx_set_grad_window": (_i, [_vp, _i]),
    "qx_get_grad_window": (_i, [_vp, _vp]),
    "qx_get_grad_wave_reduce": (_i, [_vp, _vp]),
    "qx_get_class_tiles": (_i, [_vp, _i, _vp]),
    "qx_jk_build_masked": (_i, [_vp, _vp, _vp, _vp, _i, ctypes.c_double, _vp, _vp]),
    "qx_scf_create": (_i, [_vp, _i, _i, _ip, _ip, _ip, _dp, _vp, _vp, _vp, _i, _i, _i, ctypes.c_double, _i, _i, _i,
                           ctypes.c_double, ctypes.POINTER(_vp)]),
    "qx_scf_init": (_i, [_vp, _vp, _vp]),
    "qx_scf_capture": (_i, [_vp, _vp]),
    "qx_scf_step": (_i, [_vp, _i, _i, _vp]),
    "qx_scf_status": (_i, [_vp, _dp, _dp, _dp, _ip, _ip, _dp, _vp]),
    "qx_scf_matrix": (_i, [_vp, _i, _vp, _vp]),
    "qx_scf_destroy": (None, [_vp]),
    "qx_scf_last_error": (ctypes.c_char_p, [_vp]),
    "qx_class_info": (_i, [_i, _ip, _ip]),
    "qx_class_quartets": (_i, [_vp, ctypes.POINTER(_ll), ctypes.POINTER(_ll)]),
    "qx_class_flops": (_i, [_i, _dp, _dp, _ip]),
    "qx_class_groups": (_i, [_i, _ip, _dp, _ip]),
    "qx_class_jk": (_i, [_i, _ip, _dp]),
    "qx_max_parts": (_i, []),
    "qx_profile_enable": (_i, [_vp, _i]),
    "qx_profile_read": (_i, [_vp, _dp, ctypes.POINTER(_ll), ctypes.POINTER(_ll)]),
}


def load(path: str = LIB_PATH):
    """Load the library and bind every exported symbol (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise NativeUnavailable(
            f"{path} is not built; run `python -m quax_amd.csrc.build` (hipcc, gfx950)")
    L = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int, ctx=None, what: str = ""):
    if rc != QX_OK:
        L = load()
        raw = L.qx_last_error(ctx)
        msg = raw.decode() if raw else ""
        raise QuaxNativeError(rc, f"{what}: {msg}" if what else msg)


def ptr(t):
    """Device (or host) address of a torch tensor / numpy array."""
    if hasattr(t, "data_ptr"):
        return ctypes.c_void_p(t.data_ptr())
    return ctypes.c_void_p(t.ctypes.data)


def stream_handle(device=None):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
