"""GPU parity of direct J/K, the one-electron integrals and their
derivatives against the CPU oracle."""
import os

import numpy as np
import pytest

import oracle
from quax_amd import workloads
from quax_amd.basis import BasisSet, basis_transform
from test_gpu_eri import contracted_d_cluster, small_d_cluster

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _engine(mol, basis, tol=1e-14):
    from quax_amd.engine import EriEngine
    return EriEngine(mol, basis, ints_tol=tol)


def _density(n, seed=1):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n)) * 0.3
    return X @ X.T / n


@pytest.mark.parametrize("maker", [small_d_cluster, contracted_d_cluster])
def test_direct_jk_vs_tensor(gpu, maker):
    mol, bs = maker()
    ob = oracle.OracleBasis(mol, bs)
    G = ob.compute_2e_int(1e-14)
    D = _density(ob.nbf)
    Jr = np.einsum("pqrs,rs->pq", G, D)
    Kr = np.einsum("prqs,rs->pq", G, D)
    with _engine(mol, bs) as eng:
        J, K = eng.jk(D)
        J, K = J.cpu().numpy(), K.cpu().numpy()
    assert np.abs(J - Jr).max() < 1e-11
    assert np.abs(K - Kr).max() < 1e-11


def test_direct_jk_spherical(gpu):
    mol, bs = small_d_cluster()
    bp = BasisSet(bs.basis_name, bs.shells, pure=True)
    ob = oracle.OracleBasis(mol, bs)
    T = basis_transform(bp)
    G = np.einsum("ip,jq,kr,ls,pqrs->ijkl", T, T, T, T, ob.compute_2e_int(1e-14), optimize=True)
    D = _density(bp.nbf(), 4)
    with _engine(mol, bp) as eng:
        J, K = eng.jk(D)
    assert np.abs(J.cpu().numpy() - np.einsum("pqrs,rs->pq", G, D)).max() < 1e-11
    assert np.abs(K.cpu().numpy() - np.einsum("prqs,rs->pq", G, D)).max() < 1e-11


@pytest.mark.parametrize("maker", [small_d_cluster, contracted_d_cluster])
def test_oei_vs_oracle(gpu, maker):
    mol, bs = maker()
    ob = oracle.OracleBasis(mol, bs)
    S0, T0, V0 = ob.oei()
    with _engine(mol, bs) as eng:
        S, T, V = (x.cpu().numpy() for x in eng.oei())
    for a, b in ((S, S0), (T, T0), (V, V0)):
        assert np.abs(a - b).max() < 1e-12


@pytest.mark.parametrize("maker", [small_d_cluster, contracted_d_cluster])
def test_oei_deriv_and_grad_vs_oracle(gpu, maker):
    mol, bs = maker()
    ob = oracle.OracleBasis(mol, bs)
    dS0, dT0, dV0 = ob.oei_deriv1()
    P = _density(ob.nbf, 5)
    W = _density(ob.nbf, 6)
    ref = np.einsum("xpq,pq->x", dT0 + dV0, P) + np.einsum("xpq,pq->x", dS0, W)
    with _engine(mol, bs) as eng:
        dS, dT, dV = (x.cpu().numpy() for x in eng.oei_deriv())
        g = eng.oei_grad(P, W).cpu().numpy()
    for a, b in ((dS, dS0), (dT, dT0), (dV, dV0)):
        assert np.abs(a - b).max() < 1e-11
    assert np.abs(g - ref).max() < 1e-10


@pytest.mark.parametrize("maker", [small_d_cluster, contracted_d_cluster])
def test_jk_rows_do_not_change_results(gpu, maker):
    """K' rows staged in LDS (the default) or every K' term added to global
    memory (qx_set_jk_rows off): the same J and K up to summation order."""
    from quax_amd.engine import EriEngine
    mol, basis = maker()
    rng = np.random.default_rng(3)
    n = basis.nbf()
    X = rng.standard_normal((n, n)) * 0.3
    D = X @ X.T / n
    with EriEngine(mol, basis, ints_tol=1e-14) as eng:
        nc = eng.L.qx_num_classes()
        assert eng.jk_rows().tolist() == [1] * nc
        J1, K1 = (x.cpu().numpy() for x in eng.jk(D))
        eng.set_jk_rows([0] * nc)
        assert eng.jk_rows().tolist() == [0] * nc
        J0, K0 = (x.cpu().numpy() for x in eng.jk(D))
        eng.set_jk_rows(None)
    assert np.abs(J1 - J0).max() <= 1e-13 * max(1.0, np.abs(J1).max())
    assert np.abs(K1 - K0).max() <= 1e-13 * max(1.0, np.abs(K1).max())
