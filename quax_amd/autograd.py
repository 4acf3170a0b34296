"""Differentiable two-electron energy behind torch.autograd.

    E2(R, D) = sum_pqrs (2 D_pq D_rs - D_pr D_qs) (pq|rs)
             = sum_pq D_pq (2 J - K)_pq

is the two-electron part of the RHF energy E = sum (F + H) D + E_nuc with
F = H + 2J - K (hartree_fock.py:25-28, 59, 79-82).  Its forward pass is one
integral-direct J/K build; its backward pass is the hand-written adjoint
kernel for the geometry (dE2/dR = sum (2 D D - D D) d(pq|rs)/dR, integrals
recomputed, the derivative tensor never stored) and 2(2J - K) for the
density.  The reference reaches the same gradient by forward-mode jacfwd
through the materialised derivative tensor (core.py:57-101, tei.py:25-34).
"""
from __future__ import annotations

from ._lib import QX_PREC_FP64


def _torch():
    import torch
    return torch


class _EriEnergy:
    """Holder so the autograd.Function is created only when torch imports."""
    fn = None


def _function():
    if _EriEnergy.fn is not None:
        return _EriEnergy.fn
    torch = _torch()

    class EriEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, geom, D, engine, precision, fp32_threshold):
            ctx.engine, ctx.precision, ctx.thr = engine, precision, fp32_threshold
            ctx.geom_device, ctx.d_device = geom.device, D.device
            Dd = D.detach().to(device=engine.device, dtype=torch.float64).contiguous()
            J, K = engine.jk(Dd, precision=precision, fp32_threshold=fp32_threshold)
            F2 = 2.0 * J - K
            ctx.save_for_backward(Dd, F2)
            return torch.sum(Dd * F2).to(geom.device)

        @staticmethod
        def backward(ctx, gout):
            Dd, F2 = ctx.saved_tensors
            eng = ctx.engine
            g_geom = g_D = None
            if ctx.needs_input_grad[0]:
                g = eng.eri_grad(Dd, precision=ctx.precision, fp32_threshold=ctx.thr)
                g_geom = (gout.to(g.device) * g).reshape(-1, 3).to(ctx.geom_device)
            if ctx.needs_input_grad[1]:
                g_D = (gout.to(F2.device) * 2.0 * F2).to(ctx.d_device)
            return g_geom, g_D, None, None, None

    _EriEnergy.fn = EriEnergy
    return EriEnergy


def eri_energy(engine, geom, D, precision: int = QX_PREC_FP64, fp32_threshold: float = 0.0):
    """E2(R, D) for the engine's molecule.  `geom` is an (natom, 3) tensor in
    bohr equal to the engine's geometry (the integrals are evaluated there);
    gradients flow to `geom` (adjoint kernel) and to `D` (2(2J - K))."""
    torch = _torch()
    ref = torch.as_tensor(engine.molecule.geometry(), dtype=torch.float64).reshape(-1, 3)
    if tuple(geom.shape) != tuple(ref.shape) or float((geom.detach().cpu() - ref).abs().max()) > 1e-12:
        raise ValueError("eri_energy: geom must equal the engine's geometry (build an engine per geometry)")
    return _function().apply(geom, D, engine, int(precision), float(fp32_threshold))


def jax_eri_energy(engine, precision: int = QX_PREC_FP64, fp32_threshold: float = 0.0):
    """The same E2(R, D) as a JAX function with a custom VJP that routes to
    the adjoint kernel (Quax's JAX front end; eager use, the engine holds the
    geometry).  Returns f(geom, D) -> scalar."""
    import jax  # pragma: no cover - JAX is not installed in the test image
    import jax.numpy as jnp
    import numpy as np
    torch = _torch()

    def e2_and_fock(D):
        Dd = torch.as_tensor(np.asarray(D), dtype=torch.float64, device=engine.device).contiguous()
        J, K = engine.jk(Dd, precision=precision, fp32_threshold=fp32_threshold)
        F2 = 2.0 * J - K
        return float(torch.sum(Dd * F2)), Dd, F2

    @jax.custom_vjp
    def e2(geom, D):
        return jnp.asarray(e2_and_fock(D)[0])

    def fwd(geom, D):
        E, Dd, F2 = e2_and_fock(D)
        # residuals must be JAX arrays: the density and 2J - K, on the host
        return jnp.asarray(E), (jnp.asarray(Dd.cpu().numpy()), jnp.asarray(F2.cpu().numpy()), jnp.zeros_like(geom))

    def bwd(res, gout):
        D, F2, gz = res
        Dd = torch.as_tensor(np.asarray(D), dtype=torch.float64, device=engine.device).contiguous()
        g = engine.eri_grad(Dd, precision=precision, fp32_threshold=fp32_threshold).cpu().numpy()
        return (gout * jnp.asarray(g).reshape(gz.shape), gout * 2.0 * F2)

    e2.defvjp(fwd, bwd)
    return e2
