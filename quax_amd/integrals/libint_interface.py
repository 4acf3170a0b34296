"""Drop-in replacement for Quax's pybind11 module ``libint_interface``
(quax/integrals/libint_interface.cc:2784-2807), ERI path only.

Same function names, arguments, module-global state and return conventions
(flat float64 numpy arrays the caller reshapes, tei.py:33, 118-119).  The work
runs on the GPU through libquax_hip.so; there is no CPU fallback.

Differences, all deliberate:
* ``type`` other than ``"eri"`` raises ``NotImplementedError`` for the F12
  operators (out of scope) and ``ValueError`` for unknown names (the reference
  throws std::invalid_argument -> ValueError, libint_interface.cc:520).
* A wrong ``deriv_vec`` length raises ``ValueError`` instead of aborting the
  process (libint_interface.cc:1098).
* Derivatives up to second order are built (``LIBINT2_MAX_DERIV_ORDER`` is 2
  here; the reference's value is whatever its Libint build chose).
* The four basis names must agree (the HF path, ints.py:19-20).
"""
from __future__ import annotations

import numpy as np

from ..basis import BasisSet
from ..engine import EriEngine
from ..molecule import Molecule, read_xyz

LIBINT2_MAX_DERIV_ORDER = 2
_F12_TYPES = ("f12", "f12g12", "f12_squared", "f12_double_commutator")

_engine: EriEngine | None = None
_natom = 0


def _get():
    if _engine is None:
        raise RuntimeError("libint_interface.initialize() must be called first")
    return _engine


def initialize(xyzfilename: str, basis1: str, basis2: str, basis3: str, basis4: str, ints_tol: float,
               pure: bool = False, device: int = 0) -> None:
    """libint_interface.cc:38-95: read the xyz file (Angstrom -> bohr), build
    the named basis (Cartesian unless ``pure``), set the Schwarz threshold."""
    global _engine, _natom
    if not (basis1 == basis2 == basis3 == basis4):
        raise NotImplementedError("mixed basis sets (CABS / F12) are outside the ERI path")
    symbols, coords = read_xyz(xyzfilename)
    mol = Molecule(symbols, coords)
    basis = BasisSet.build(mol, basis1, pure=pure)
    finalize()
    _engine = EriEngine(mol, basis, ints_tol=ints_tol, device=device)
    _natom = mol.natom()


def finalize() -> None:
    """libint_interface.cc:97-99."""
    global _engine
    if _engine is not None:
        _engine.close()
        _engine = None


def _check_type(type_: str):
    if type_ == "eri":
        return
    if type_ in _F12_TYPES:
        raise NotImplementedError(f"operator '{type_}' is outside the ERI path")
    raise ValueError("type must be eri, f12, f12g12, f12_squared, or f12_double_commutator")


def compute_2e_int(type_: str, beta: float) -> np.ndarray:
    """libint_interface.cc:486-647: flat nbf^4 array, chemists' notation."""
    _check_type(type_)
    G = _get().compute_2e_int()
    return G.reshape(-1).cpu().numpy()


def eri_deriv_core(deriv_order: int) -> np.ndarray:
    """libint_interface.cc:2582-2779: flat (how_many_derivs * nbf^4)."""
    if deriv_order > LIBINT2_MAX_DERIV_ORDER or deriv_order < 1:
        raise NotImplementedError(f"eri_deriv_core({deriv_order}): orders 1 to {LIBINT2_MAX_DERIV_ORDER} are built")
    return _get().eri_deriv_core(deriv_order).reshape(-1).cpu().numpy()


def compute_2e_deriv(type_: str, beta: float, deriv_vec) -> np.ndarray:
    """libint_interface.cc:1097-1335: one flat nbf^4 derivative slice."""
    _check_type(type_)
    dv = np.asarray(deriv_vec, dtype=np.int64).reshape(-1)
    if dv.shape[0] != 3 * _natom:
        raise ValueError("Derivative vector incorrect size for this molecule.")
    if int(dv.sum()) > LIBINT2_MAX_DERIV_ORDER:
        raise NotImplementedError(f"compute_2e_deriv: deriv_vec orders up to {LIBINT2_MAX_DERIV_ORDER} are built")
    return _get().compute_2e_deriv(dv).reshape(-1).cpu().numpy()


def compute_2e_deriv_disk(type_: str, beta: float, max_deriv_order: int) -> None:
    """libint_interface.cc:1811-2115: writes every derivative tensor of
    orders 1..max_deriv_order to <type>_derivs.h5 in the working directory,
    one dataset "<type>_<n>_<n>_<n>_<n>_deriv<order>" of shape (n, n, n, n,
    how_many_derivs) per order (derivative index last, zero fill).  The
    tensors come from eri_deriv_core on the GPU, transposed there and written
    one first-index row at a time (libquax_h5.so).  Like the reference, a
    request above 50 GB of tensors is refused (ValueError here, an assert
    there)."""
    from .. import _h5
    _check_type(type_)
    if max_deriv_order < 1 or max_deriv_order > LIBINT2_MAX_DERIV_ORDER:
        raise NotImplementedError(f"compute_2e_deriv_disk: orders 1 to {LIBINT2_MAX_DERIV_ORDER} are built")
    eng = _get()
    n = eng.nbf
    total = sum(how_many_derivs(_natom, k) for k in range(1, max_deriv_order + 1))
    if n ** 4 * total * 8 * 1e-9 >= 50:
        raise ValueError("Total disk space required for ERI's exceeds 50 GB.")
    path = f"{type_}_derivs.h5"
    mode = _h5.TRUNCATE
    for order in range(1, max_deriv_order + 1):
        nd = how_many_derivs(_natom, order)
        T = eng.eri_deriv_core(order).reshape(nd, n, n, n, n)
        name = f"{type_}_{n}_{n}_{n}_{n}_deriv{order}"
        for i in range(n):
            row = T[:, i].permute(1, 2, 3, 0).contiguous().cpu().numpy()
            _h5.write_block(path, name, (n, n, n, n, nd), i, row[None], mode=mode)
            mode = _h5.APPEND
        del T


def generate_multi_index_lookup(nparams: int, deriv_order: int):
    """utils.h:142-172: combinations with replacement, lexicographic."""
    import ctypes
    L = _get_lib()
    n = L.qx_generate_multi_index_lookup(nparams, deriv_order, None, 0)
    if n <= 0:
        return [[]] if n == 1 else []
    buf = np.zeros(max(n * deriv_order, 1), dtype=np.int32)
    L.qx_generate_multi_index_lookup(nparams, deriv_order, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n)
    return buf[: n * deriv_order].reshape(n, deriv_order).tolist()


def how_many_derivs(natom: int, deriv_order: int) -> int:
    """utils.h:131-140."""
    return int(_get_lib().qx_how_many_derivs(natom, deriv_order))


def schwarz_screening():
    """libint_interface.cc:158-229: (pair list, nshell x nshell values)."""
    vals, pairs = _get().schwarz_screening()
    return [tuple(p) for p in pairs.tolist()], vals


def engine() -> EriEngine:
    return _get()


def _get_lib():
    from .. import _lib
    return _lib.load()
