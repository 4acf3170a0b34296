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
        if mode == "disk":
            raise NotImplementedError("the HDF5 disk mode is outside the ERI path")
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
        G = libint_interface.compute_2e_deriv("eri", 0.0, deriv_vec)
        return _asarray(G.reshape(self.nbf1, self.nbf2, self.nbf3, self.nbf4))

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
