"""User-facing HF energy and nuclear gradient on the GPU ERI path.

Mirrors the HF branch of quax.core.energy / geom_deriv (core.py:302-393):
same molecule, basis-name and options arguments, the gradient rounded to
10 decimals (core.py:88).  The reference differentiates forward through the
SCF with jacfwd over a precomputed (3N, n, n, n, n) derivative tensor; here
the gradient is analytic, with the two-electron part from the adjoint kernel
that recomputes derivative integrals and contracts them with the density on
the fly.
"""
from __future__ import annotations

import numpy as np

from .basis import BasisSet
from .engine import EriEngine
from .methods.hartree_fock import check_options, restricted_hartree_fock, rhf_gradient, rhf_hessian_vector_product


def _setup(molecule, basis_name, options):
    opts = check_options(options)
    basis = BasisSet.build(molecule, basis_name, pure=bool(opts["pure"]))
    eng = EriEngine(molecule, basis, ints_tol=opts["ints_tolerance"], device=opts["device"])
    nelec = molecule.nelectron()
    if nelec % 2:
        raise ValueError("RHF needs an even electron count")
    return opts, eng, nelec


def energy(molecule, basis_name, method="hf", options=None, return_aux_data=False):
    if method not in ("hf", "scf", "rhf"):
        raise NotImplementedError(f"method '{method}' is outside the ERI path")
    opts, eng, nelec = _setup(molecule, basis_name, options)
    with eng:
        E, aux = restricted_hartree_fock(eng, nelec, opts, return_aux_data=True)
    return (E, aux) if return_aux_data else E


def geom_deriv(molecule, basis_name, method="hf", deriv_order=1, partial=None, options=None):
    """Gradient (deriv_order 1) or Hessian (deriv_order 2, rounded like
    core.py:89-91), the Hessian one column per nuclear coordinate as
    forward-over-reverse Hessian-vector products (rhf_hessian_vector_product)."""
    if method not in ("hf", "scf", "rhf"):
        raise NotImplementedError(f"method '{method}' is outside the ERI path")
    if deriv_order not in (1, 2):
        raise NotImplementedError("first and second derivatives are built")
    if partial is not None and len(partial) != deriv_order:
        raise Exception("The length of the index coordinates given by 'partial' argument should be the same "
                        "as the order of differentiation")
    opts, eng, nelec = _setup(molecule, basis_name, options)
    with eng:
        E, aux = restricted_hartree_fock(eng, nelec, opts, return_aux_data=True)
        if deriv_order == 1:
            g = rhf_gradient(eng, aux, precision=opts["precision"], fp32_threshold=opts["fp32_threshold"])
            return np.round(g[partial[0]], 14) if partial is not None else np.round(g, 10)
        m = 3 * molecule.natom()
        cols = [partial[1]] if partial is not None else range(m)
        hess = np.zeros((m, m))
        for k in cols:
            e = np.zeros(m)
            e[k] = 1.0
            hess[:, k] = rhf_hessian_vector_product(eng, aux, e)
    if partial is not None:
        return np.round(hess[partial[0], partial[1]], 14)
    return np.round(hess, 10)
