"""ctypes binding of libquax_h5.so (include/quax_amd_h5.h): HDF5 files of
integral derivative tensors in the layout of Quax's disk mode.  Raises
NativeUnavailable if the library was not built (no HDF5 on the build host)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._lib import NativeUnavailable

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libquax_h5.so")
READ, TRUNCATE, APPEND = 0, 1, 2
MAX_RANK = 8

_lib = None
_ll = ctypes.c_longlong
_SIG = {
    "qx_h5_write_block": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                         ctypes.POINTER(_ll), _ll, _ll, ctypes.c_void_p]),
    "qx_h5_dataset_dims": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(_ll)]),
    "qx_h5_read_slice": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _ll, ctypes.c_void_p, _ll]),
    "qx_h5_last_error": (ctypes.c_char_p, []),
}


class H5Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[quax_amd h5 code {code}] {msg}")
        self.code = code


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeUnavailable(f"{LIB_PATH} is not built (HDF5 headers and library under /opt/conda needed)")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIG.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise H5Error(rc, (load().qx_h5_last_error() or b"").decode())


def write_block(path, dset, dims, first, block, mode=APPEND):
    """Rows [first, first + len(block)) along dimension 0 of dataset dset of
    shape dims (created zero-filled if absent)."""
    L = load()
    b = np.ascontiguousarray(block, dtype=np.float64)
    d = (_ll * len(dims))(*[int(x) for x in dims])
    if b.shape[1:] != tuple(int(x) for x in dims[1:]):
        raise ValueError("block rows must have the dataset's trailing shape")
    _check(L.qx_h5_write_block(os.fsencode(path), int(mode), dset.encode(), len(dims), d, int(first), b.shape[0],
                               b.ctypes.data_as(ctypes.c_void_p)))


def dataset_dims(path, dset):
    """Shape of a dataset, or None if the file or dataset does not exist."""
    L = load()
    r = ctypes.c_int()
    d = (_ll * MAX_RANK)()
    rc = L.qx_h5_dataset_dims(os.fsencode(path), dset.encode(), ctypes.byref(r), d)
    if rc in (2, 4):
        return None
    _check(rc)
    return tuple(d[k] for k in range(r.value))


def read_slice(path, dset, index=0):
    """Rank-5 dataset: [:, :, :, :, index]; otherwise the whole dataset."""
    dims = dataset_dims(path, dset)
    if dims is None:
        raise H5Error(4, f"{path}: no dataset {dset}")
    shape = dims[:4] if len(dims) == 5 else dims
    out = np.empty(shape, dtype=np.float64)
    _check(load().qx_h5_read_slice(os.fsencode(path), dset.encode(), int(index), out.ctypes.data_as(ctypes.c_void_p),
                                   out.size))
    return out
