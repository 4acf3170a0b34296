"""Two-electron integral primitive layer, mirroring quax/integrals/tei.py for
the Coulomb ERI (tei.py:12-44, 86-90, 117-120, 142-155, 336-346, 403-411).

The reference binds ``jax.core.Primitive`` objects whose ``impl`` calls the
native module, whose JVP maps d eri[t] -> eri_deriv(t) and
d eri_deriv(dv)[t] -> eri_deriv(dv + t), and whose batching rule stacks
``eri_deriv`` over the rows of a batched ``deriv_vec``.  This class keeps
those rules verbatim as plain methods (usable without JAX) and, when JAX is
importable, registers them on real primitives the same way, adding an
abstract-eval rule so the primitives also work under ``jit``.
"""
from __future__ import annotations

import os

import numpy as np

from . import libint_interface
from ..molecule import read_xyz
from ..utils import get_deriv_vec_idx, how_many_derivs

try:  # JAX is optional; the primitive location moved in later releases
    import jax  # noqa: F401
    import jax.numpy as jnp
    try:
        from jax.extend.core import Primitive as _Primitive
    except Exception:  # pragma: no cover - older JAX
        from jax.core import Primitive as _Primitive
    _HAVE_JAX = True
except Exception:
    jnp = None
    _Primitive = None
    _HAVE_JAX = False


def _asarray(x):
    return jnp.asarray(x) if _HAVE_JAX else np.asarray(x)


class TEI:
    def __init__(self, basis1, basis2, basis3, basis4, xyz_path, max_deriv_order, options, mode):
        symbols, _ = read_xyz(xyz_path)
        natoms = len(symbols)
        self.natoms = natoms
        nbf1, nbf2, nbf3, nbf4 = basis1.nbf(), basis2.nbf(), basis3.nbf(), basis4.nbf()
        if mode not in ("core", "f12", "disk"):
            raise ValueError(f"unknown TEI mode '{mode}'")
        if mode == "core" and max_deriv_order > 0:
            self.eri_derivatives = []
            for i in range(max_deriv_order):
                n_unique = how_many_derivs(natoms, i + 1)
                d = libint_interface.eri_deriv_core(i + 1).reshape(n_unique, nbf1, nbf2, nbf3, nbf4)
                self.eri_derivatives.append(d)
        self.mode = mode
        self.nbf1, self.nbf2, self.nbf3, self.nbf4 = nbf1, nbf2, nbf3, nbf4
        if _HAVE_JAX:
            self._register_jax()

    # -- calls (tei.py:86-90) --
    def eri(self, geom):
        if _HAVE_JAX:
            return self.eri_p.bind(geom)
        return self.eri_impl(geom)

    def eri_deriv(self, geom, deriv_vec):
        if _HAVE_JAX:
            return self.eri_deriv_p.bind(geom, deriv_vec)
        return self.eri_deriv_impl(geom, deriv_vec)

    # -- evaluation rules (tei.py:117-120, 142-155) --
    def eri_impl(self, geom):
        G = libint_interface.compute_2e_int("eri", 0.0)
        return _asarray(G.reshape(self.nbf1, self.nbf2, self.nbf3, self.nbf4))

    def eri_deriv_impl(self, geom, deriv_vec):
        deriv_vec = np.asarray(deriv_vec, int)
        deriv_order = int(np.sum(deriv_vec))
        idx = get_deriv_vec_idx(deriv_vec)
        if self.mode == "core":
            return _asarray(self.eri_derivatives[deriv_order - 1][idx, :, :, :, :])
        if self.mode == "disk":
            return _asarray(self._read_disk(deriv_order, idx))
        G = libint_interface.compute_2e_deriv("eri", 0.0, deriv_vec)
        return _asarray(G.reshape(self.nbf1, self.nbf2, self.nbf3, self.nbf4))

    def _read_disk(self, deriv_order, idx):
        """tei.py:156-181: the full tensor file first, then the partials file."""
        from .. import _h5
        dims = f"{self.nbf1}_{self.nbf2}_{self.nbf3}_{self.nbf4}"
        if os.path.exists("eri_derivs.h5"):
            file_name, dataset_name = "eri_derivs.h5", f"eri_{dims}_deriv{deriv_order}"
        elif os.path.exists("eri_partials.h5"):
            file_name, dataset_name = "eri_partials.h5", f"eri_{dims}_deriv{deriv_order}_{idx}"
        else:
            raise Exception("ERI derivatives not found on disk")
        shape = _h5.dataset_dims(file_name, dataset_name)
        if shape is None or len(shape) not in (4, 5):
            raise Exception("Something went wrong reading integral derivative file")
        return _h5.read_slice(file_name, dataset_name, idx)

    # -- JVP rules (tei.py:336-346) --
    def eri_jvp(self, primals, tangents):
        geom, = primals
        return self.eri(geom), self.eri_deriv(geom, tangents[0])

    def eri_deriv_jvp(self, primals, tangents):
        geom, deriv_vec = primals
        return self.eri_deriv(geom, deriv_vec), self.eri_deriv(geom, deriv_vec + tangents[0])

    # -- batching rule (tei.py:403-411) --
    def eri_deriv_batch(self, batched_args, batch_dims):
        geom_batch, deriv_batch = batched_args
        results = [self.eri_deriv(geom_batch, i)[None] for i in deriv_batch]
        if _HAVE_JAX:
            return jnp.concatenate(results, axis=0), 0
        return np.concatenate(results, axis=0), 0

    # -- JAX registration (tei.py:43-83) --
    def _register_jax(self):  # pragma: no cover - JAX absent in the test image
        from jax.interpreters import ad, batching
        import jax.core as jcore
        self.eri_p = _Primitive("eri")
        self.eri_deriv_p = _Primitive("eri_deriv")
        self.eri_p.def_impl(self.eri_impl)
        self.eri_deriv_p.def_impl(self.eri_deriv_impl)
        shape = (self.nbf1, self.nbf2, self.nbf3, self.nbf4)
        aval = jcore.ShapedArray(shape, jnp.float64)
        self.eri_p.def_abstract_eval(lambda geom: aval)
        self.eri_deriv_p.def_abstract_eval(lambda geom, dv: aval)
        ad.primitive_jvps[self.eri_p] = self.eri_jvp
        ad.primitive_jvps[self.eri_deriv_p] = self.eri_deriv_jvp
        batching.primitive_batchers[self.eri_deriv_p] = self.eri_deriv_batch


def check_tei_disk(int_type, basis1, basis2, basis3, basis4, deriv_order, address=None):
    """ints.py:331-360: True when <int_type>_derivs.h5 in the working
    directory holds the order-`deriv_order` dataset with these basis sizes.
    (The reference scans the file's keys for names contained in the wanted
    name; the exact name is looked up here.)"""
    from .. import _h5
    path = int_type + "_derivs.h5"
    if not os.path.exists(path):
        return False
    nbf = (basis1.nbf(), basis2.nbf(), basis3.nbf(), basis4.nbf())
    name = int_type + "_" + "_".join(str(x) for x in nbf) + "_deriv" + str(deriv_order)
    dims = _h5.dataset_dims(path, name)
    ok = dims is not None and tuple(dims[:4]) == nbf
    if ok:
        print("Integral derivatives appear to be correct. Avoiding recomputation.")
    return ok
