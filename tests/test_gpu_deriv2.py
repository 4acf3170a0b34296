"""Second derivatives and Hessian-vector products on the GPU.

* eri_deriv_core(2) and compute_2e_deriv at orders 0 and 2 (tangent,
  forward-mode kernels over the first-derivative recurrences) against the CPU
  oracle's restatement of libint_interface.cc:1097-1335 / 2582-2779, 1e-12
  per element.
* The forward-over-reverse pieces: the adjoint contraction of second
  derivatives along a direction v against the oracle's second-derivative
  tensor contracted on the host; the derivative J/K along v against the
  oracle's first-derivative tensor; the one-electron second-derivative
  contraction against central differences of the oracle's one-electron
  derivative tensors.
* The RHF Hessian-vector product (CPHF response + the pieces above) against
  Richardson-extrapolated central differences of analytic gradients: water
  STO-3G (the differentiation molecule) and the 4-atom cluster subset with
  the 112-function spherical basis.
"""
import numpy as np
import pytest

import oracle
from quax_amd import workloads
from quax_amd.basis import BasisSet
from quax_amd.molecule import Molecule
from test_gpu_eri import contracted_d_cluster

pytestmark = pytest.mark.gpu
ATOL = 1e-12


def _engine(mol, basis, tol=1e-14):
    from quax_amd.engine import EriEngine
    return EriEngine(mol, basis, ints_tol=tol)


def _water():
    mol = workloads.water_sto3g()
    return mol, BasisSet.build(mol, "sto-3g")


def _pairs(m):
    return [(x, y) for x in range(m) for y in range(x, m)]


def _density(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n)) * 0.3
    return X @ X.T / n


@pytest.mark.parametrize("maker", [_water, contracted_d_cluster])
def test_deriv_core2_vs_oracle(gpu, maker):
    mol, bs = maker()
    ob = oracle.OracleBasis(mol, bs)
    ref = ob.eri_deriv_core2(1e-14)
    with _engine(mol, bs) as eng:
        d2 = eng.eri_deriv_core(2).cpu().numpy()
    m = 3 * mol.natom()
    assert d2.shape == (m * (m + 1) // 2,) + (ob.nbf,) * 4
    err = np.abs(d2 - ref).max()
    print(f"  eri_deriv_core(2) {mol.natom()} atoms: max |d| {err:.2e} (max |ref| {np.abs(ref).max():.2e})")
    assert err < ATOL


def test_compute_2e_deriv_order2_and_order0(gpu):
    mol, bs = contracted_d_cluster()
    ob = oracle.OracleBasis(mol, bs)
    m = 3 * mol.natom()
    with _engine(mol, bs) as eng:
        for (x, y) in ((1, 1), (0, 2), (2, 4), (3, 5)):
            dv = np.zeros(m, dtype=int)
            dv[x] += 1
            dv[y] += 1
            ref = ob.compute_2e_deriv2(1e-14, dv)
            d = eng.compute_2e_deriv(dv).cpu().numpy()
            assert np.abs(d - ref).max() < ATOL, (x, y)
        # order 0: the integrals with quartets on one atom left out
        d0 = eng.compute_2e_deriv(np.zeros(m, dtype=int)).cpu().numpy()
    G = ob.compute_2e_int(1e-14)
    at = np.repeat([s.atom for s in bs.shells], [(s.l + 1) * (s.l + 2) // 2 for s in bs.shells])
    single = (at[:, None, None, None] == at[None, :, None, None]) & (at[:, None, None, None] == at[None, None, :, None]) \
        & (at[:, None, None, None] == at[None, None, None, :])
    assert np.abs(d0 - np.where(single, 0.0, G)).max() < ATOL
    assert np.abs(d0[single]).max() == 0.0


def test_spherical_deriv_core2_is_transformed_cartesian(gpu):
    from quax_amd.basis import basis_transform
    mol, bs = contracted_d_cluster()
    bp = BasisSet(bs.basis_name, bs.shells, pure=True)
    T = basis_transform(bp)
    with _engine(mol, bs) as eng:
        dc = eng.eri_deriv_core(2).cpu().numpy()
    with _engine(mol, bp) as eng:
        dp = eng.eri_deriv_core(2).cpu().numpy()
    ref = np.einsum("ap,bq,cr,ds,xpqrs->xabcd", T, T, T, T, dc, optimize=True)
    assert np.abs(dp - ref).max() < ATOL


@pytest.mark.parametrize("maker", [_water, contracted_d_cluster])
def test_eri_hess_dir_vs_oracle_tensor(gpu, maker):
    mol, bs = maker()
    ob = oracle.OracleBasis(mol, bs)
    d2 = ob.eri_deriv_core2(1e-14)
    m = 3 * mol.natom()
    n = ob.nbf
    D = _density(n, 4)
    v = np.random.default_rng(9).standard_normal(m)
    # (m, m) matrix of second-derivative slices from the upper triangle
    idx = np.zeros((m, m), dtype=int)
    for q, (x, y) in enumerate(_pairs(m)):
        idx[x, y] = idx[y, x] = q
    gam = 2 * np.einsum("pq,rs->pqrs", D, D) - np.einsum("pr,qs->pqrs", D, D)
    per_slice = np.einsum("xpqrs,pqrs->x", d2, gam, optimize=True)
    ref = per_slice[idx] @ v
    with _engine(mol, bs) as eng:
        h = eng.eri_hess_dir(D, v).cpu().numpy()
    assert np.abs(h - ref).max() < 1e-10 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("maker", [_water, contracted_d_cluster])
def test_jk_deriv_dir_vs_oracle_tensor(gpu, maker):
    mol, bs = maker()
    ob = oracle.OracleBasis(mol, bs)
    d1 = ob.eri_deriv_core1(1e-14)
    n = ob.nbf
    D = _density(n, 6)
    v = np.random.default_rng(2).standard_normal(3 * mol.natom())
    dG = np.tensordot(v, d1, axes=1)
    Jr = np.einsum("pqrs,rs->pq", dG, D)
    Kr = np.einsum("prqs,rs->pq", dG, D)
    with _engine(mol, bs) as eng:
        J, K = (x.cpu().numpy() for x in eng.jk_deriv_dir(D, v))
    assert np.abs(J - Jr).max() < 1e-11
    assert np.abs(K - Kr).max() < 1e-11


@pytest.mark.parametrize("maker", [_water, contracted_d_cluster])
def test_oei_hess_dir_vs_oracle_differences(gpu, maker):
    """sum Pd d/dv d(T+V) + Wd d/dv dS against central differences (h = 1e-4,
    Richardson-extrapolated with h/2) of the oracle's derivative tensors."""
    mol, bs = maker()
    n = oracle.OracleBasis(mol, bs).nbf
    m = 3 * mol.natom()
    Pd, Wd = _density(n, 7), _density(n, 8)
    v = np.random.default_rng(5).standard_normal(m)
    xyz0 = np.asarray(mol.geometry(), dtype=np.float64).reshape(-1)

    def contr(xyz):
        mm = Molecule([mol.symbol(i) for i in range(mol.natom())], xyz.reshape(-1, 3))
        dS, dT, dV = oracle.OracleBasis(mm, bs).oei_deriv1()
        return np.einsum("xpq,pq->x", dT + dV, Pd) + np.einsum("xpq,pq->x", dS, Wd)

    def cd(h):
        return (contr(xyz0 + h * v) - contr(xyz0 - h * v)) / (2 * h)

    ref = (4 * cd(5e-5) - cd(1e-4)) / 3
    with _engine(mol, bs) as eng:
        got = eng.oei_hess_dir(Pd, Wd, v).cpu().numpy()
    assert np.abs(got - ref).max() < 1e-8 * max(1.0, np.abs(ref).max())


def _rhf(mol, bs, opts, D0=None, tol=1e-14):
    from quax_amd.engine import EriEngine
    from quax_amd.methods.hartree_fock import restricted_hartree_fock
    eng = EriEngine(mol, bs, ints_tol=tol)
    E, aux = restricted_hartree_fock(eng, mol.nelectron(), dict(opts), return_aux_data=True, D0=D0)
    assert aux["drms"] <= 1e-11, aux["drms"]
    return eng, E, aux


def _hvp_vs_fd(mol, bs, vs, opts, D0=None, tol=1e-14, h=2e-4):
    """max |H v - FD| over the directions, FD = Richardson extrapolation of
    central differences of analytic gradients with steps h and h/2 along the
    unit direction (truncation O(h^4))."""
    from quax_amd.methods.hartree_fock import rhf_gradient, rhf_hessian_vector_product
    eng, E, aux = _rhf(mol, bs, opts, D0, tol)
    vs = [np.asarray(v, dtype=np.float64) / np.linalg.norm(v) for v in vs]
    with eng:
        Hv = [rhf_hessian_vector_product(eng, aux, v) for v in vs]
        D_base = aux["D"].clone()
    xyz0 = np.asarray(mol.geometry(), dtype=np.float64).reshape(-1)

    def grad_at(x):
        mm = Molecule([mol.symbol(i) for i in range(mol.natom())], x.reshape(-1, 3))
        e1, _, a1 = _rhf(mm, bs, opts, D_base, tol)
        with e1:
            return rhf_gradient(e1, a1)

    worst = 0.0
    for v, hv in zip(vs, Hv):
        cd = [(grad_at(xyz0 + s * v) - grad_at(xyz0 - s * v)) / (2 * s) for s in (h, h / 2)]
        fd = (4 * cd[1] - cd[0]) / 3
        err = np.abs(hv - fd).max()
        print(f"  HVP {mol.natom()} atoms: max |Hv - FD| {err:.2e} (max |Hv| {np.abs(hv).max():.2e})", flush=True)
        worst = max(worst, err)
        # translational invariance: the atom blocks of H v sum to zero
        assert np.abs(hv.reshape(-1, 3).sum(0)).max() < 1e-8
    return worst


def test_hvp_water_vs_gradient_differences(gpu):
    mol, bs = _water()
    rng = np.random.default_rng(17)
    vs = [rng.standard_normal(9) for _ in range(2)] + [np.eye(9)[4]]
    opts = {"spectral_shift": False, "diis": True, "conv": 1e-12, "maxit": 200}
    assert _hvp_vs_fd(mol, bs, vs, opts) < 1e-6


def test_hvp_subset4_vs_gradient_differences(gpu):
    from test_gpu_configs import core_density
    mol = workloads.cluster10_subset4()
    bs = BasisSet.build(mol, "even-tempered", pure=True)
    opts = {"spectral_shift": False, "diis": True, "conv": 1e-12, "maxit": 300}
    with _engine(mol, bs) as eng:
        D0 = core_density(eng, mol.nelectron() // 2)
    rng = np.random.default_rng(23)
    vs = [rng.standard_normal(12)]
    assert _hvp_vs_fd(mol, bs, vs, opts, D0=D0) < 1e-6


def test_hessian_from_geom_deriv(gpu):
    """geom_deriv(deriv_order=2): symmetric, translationally invariant, and
    H[:, k] equal to the Hessian-vector product along e_k."""
    from quax_amd import core
    mol, _ = _water()
    opts = {"spectral_shift": False, "diis": True, "conv": 1e-12, "maxit": 200}
    H = core.geom_deriv(mol, "sto-3g", "hf", deriv_order=2, options=opts)
    assert H.shape == (9, 9)
    assert np.abs(H - H.T).max() < 1e-8
    assert np.abs(H.reshape(3, 3, 9).sum(0)).max() < 1e-8
    h47 = core.geom_deriv(mol, "sto-3g", "hf", deriv_order=2, partial=(4, 7), options=opts)
    assert abs(h47 - H[4, 7]) < 1e-9
