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
from .methods.hartree_fock import check_options, restricted_hartree_fock, rhf_gradient


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
    if method not in ("hf", "scf", "rhf"):
        raise NotImplementedError(f"method '{method}' is outside the ERI path")
    if deriv_order != 1:
        raise NotImplementedError("only first derivatives are built")
    opts, eng, nelec = _setup(molecule, basis_name, options)
    with eng:
        E, aux = restricted_hartree_fock(eng, nelec, opts, return_aux_data=True)
        g = rhf_gradient(eng, aux, precision=opts["precision"], fp32_threshold=opts["fp32_threshold"])
    if partial is not None:
        if len(partial) != deriv_order:
            raise Exception("The length of the index coordinates given by 'partial' argument should be the same "
                            "as the order of differentiation")
        return np.round(g[partial[0]], 14)
    return np.round(g, 10)
