"""Mixed precision and multi-stream passes on the GPU.

Quartets whose Schwarz bound sqrt(Q_ab Q_cd) is below the threshold run the
recursion in FP32 and are summed in FP64; the rest are the FP64 kernels.
Each class list is evaluated by two launches masked to its FP64 and FP32
kets, in the classes the per-class FP32 policy admits; the tests that check
the FP32 kernels admit every class.
The FP64 GPU path is the reference here: it is checked against the oracle
by the other GPU tests.  An FP32 integral carries a relative error of a few
FP32 ulps (~1e-7 of the quartet's magnitude, which the bound caps), so the
bars below are: everything in FP32 -> 1e-5 absolute on J, K and the
gradient of these O(1-10) quantities; threshold 1e-6 -> 1e-9.
"""
import numpy as np
import pytest

import oracle
from quax_amd import workloads
from quax_amd.basis import BasisSet
from quax_amd.molecule import Molecule
from test_gpu_eri import contracted_d_cluster, small_d_cluster

pytestmark = pytest.mark.gpu
MIXED = 1


def _engine(mol, basis, tol=1e-12):
    from quax_amd.engine import EriEngine
    return EriEngine(mol, basis, ints_tol=tol)


def _density(n, seed=1):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n)) * 0.3
    return X @ X.T / n


def _all_fp32(eng):
    n = len(eng.fp32_classes(0))
    for p in (0, 1):
        eng.set_fp32_classes(p, np.ones(n, dtype=bool))


def _c2():
    m = workloads.cluster10()
    return Molecule([m.symbol(i) for i in range(3)], m.geometry()[:3])


@pytest.mark.parametrize("maker", [small_d_cluster, contracted_d_cluster])
def test_split_counts_partition_the_quartets(gpu, maker):
    mol, bs = maker()
    with _engine(mol, bs) as eng:
        nq = eng.num_quartets()
        assert eng.mixed_quartets(0.0) == (nq, 0)
        assert eng.mixed_quartets(1e300) == (0, nq)
        last32 = 0
        for thr in (1e-8, 1e-6, 1e-4, 1e-2):
            n64, n32 = eng.mixed_quartets(thr)
            assert n64 + n32 == nq
            assert n32 >= last32          # a higher threshold moves more quartets to FP32
            last32 = n32
            # a pass counts FP32 quartets in the classes its policy admits only
            for p in (0, 1):
                a, b = eng.mixed_quartets(thr, p)
                assert a + b == nq and b <= n32


def test_fp32_class_policy_round_trip(gpu):
    mol, bs = small_d_cluster()
    with _engine(mol, bs) as eng:
        nq = eng.num_quartets()
        default = [eng.fp32_classes(p) for p in (0, 1)]
        # the built-in policy admits FP32 only in some (the heaviest) classes
        assert all(0 < d.sum() < len(d) for d in default)
        _all_fp32(eng)
        assert all(eng.fp32_classes(p).all() for p in (0, 1))
        assert eng.mixed_quartets(1e300, 0) == (0, nq)
        eng.set_fp32_classes(1, np.zeros(len(default[1]), dtype=bool))
        assert eng.mixed_quartets(1e300, 1) == (nq, 0)
        for p in (0, 1):
            eng.set_fp32_classes(p, None)
            assert (eng.fp32_classes(p) == default[p]).all()
        with pytest.raises(ValueError):
            eng.set_fp32_classes(0, [1, 0])


@pytest.mark.parametrize("maker", [small_d_cluster, contracted_d_cluster])
def test_mixed_jk_and_gradient_vs_fp64(gpu, maker):
    mol, bs = maker()
    with _engine(mol, bs) as eng:
        _all_fp32(eng)
        D = _density(eng.nbf, 3)
        J, K = (x.cpu().numpy() for x in eng.jk(D))
        g = eng.eri_grad(D).cpu().numpy()
        # threshold 0: every quartet in FP64 -> the FP64 result
        J0, K0 = (x.cpu().numpy() for x in eng.jk(D, precision=MIXED, fp32_threshold=0.0))
        assert np.abs(J0 - J).max() < 1e-12 and np.abs(K0 - K).max() < 1e-12
        # everything in FP32
        J32, K32 = (x.cpu().numpy() for x in eng.jk(D, precision=MIXED, fp32_threshold=1e300))
        g32 = eng.eri_grad(D, precision=MIXED, fp32_threshold=1e300).cpu().numpy()
        assert 0 < np.abs(J32 - J).max() < 1e-5
        assert 0 < np.abs(K32 - K).max() < 1e-5
        assert np.abs(g32 - g).max() < 1e-5
        # a realistic threshold: only small quartets in FP32
        Jm, Km = (x.cpu().numpy() for x in eng.jk(D, precision=MIXED, fp32_threshold=1e-6))
        gm = eng.eri_grad(D, precision=MIXED, fp32_threshold=1e-6).cpu().numpy()
        assert np.abs(Jm - J).max() < 1e-9
        assert np.abs(Km - K).max() < 1e-9
        assert np.abs(gm - g).max() < 1e-9
        # a threshold that splits most bra rows between the two masked
        # launches: a quartet lost or counted twice would move J, K or the
        # gradient by the size of an integral
        thr = next(t for t in np.logspace(-4, 1, 16)
                   if all(min(eng.mixed_quartets(t, p)) > 0 for p in (0, 1)))
        Js, Ks = (x.cpu().numpy() for x in eng.jk(D, precision=MIXED, fp32_threshold=thr))
        gs = eng.eri_grad(D, precision=MIXED, fp32_threshold=thr).cpu().numpy()
        assert np.abs(Js - J).max() < 1e-5
        assert np.abs(Ks - K).max() < 1e-5
        assert np.abs(gs - g).max() < 1e-5
        # the built-in policy: FP64 kernels in the classes it does not admit
        for p in (0, 1):
            eng.set_fp32_classes(p, None)
        Jd, Kd = (x.cpu().numpy() for x in eng.jk(D, precision=MIXED, fp32_threshold=1e300))
        gd = eng.eri_grad(D, precision=MIXED, fp32_threshold=1e300).cpu().numpy()
        assert np.abs(Jd - J).max() < 1e-5 and np.abs(Kd - K).max() < 1e-5
        assert np.abs(gd - g).max() < 1e-5


def test_mixed_fp32_matches_oracle_loosely(gpu):
    """FP32 kernels of every class against the oracle's tensor directly."""
    mol, bs = small_d_cluster()
    ob = oracle.OracleBasis(mol, bs)
    G = ob.compute_2e_int(1e-14)
    D = _density(ob.nbf, 5)
    with _engine(mol, bs, tol=1e-14) as eng:
        _all_fp32(eng)
        J, K = (x.cpu().numpy() for x in eng.jk(D, precision=MIXED, fp32_threshold=1e300))
    assert np.abs(J - np.einsum("pqrs,rs->pq", G, D)).max() < 1e-5
    assert np.abs(K - np.einsum("prqs,rs->pq", G, D)).max() < 1e-5


def test_streams_do_not_change_results(gpu):
    mol = _c2()
    bs = BasisSet.build(mol, "even-tempered", pure=True)
    with _engine(mol, bs) as eng:
        D = _density(eng.nbf, 2)
        eng.set_streams(1)
        J1, K1 = (x.cpu().numpy() for x in eng.jk(D))
        g1 = eng.eri_grad(D).cpu().numpy()
        eng.set_streams(4)
        J4, K4 = (x.cpu().numpy() for x in eng.jk(D))
        g4 = eng.eri_grad(D).cpu().numpy()
    # only the order of FP64 atomic additions differs
    assert np.abs(J4 - J1).max() < 1e-12
    assert np.abs(K4 - K1).max() < 1e-12
    assert np.abs(g4 - g1).max() < 1e-12


def test_mixed_rhf_energy_and_gradient(gpu):
    """RHF on a 3-atom even-tempered cluster: mixed at 1e-8 vs FP64."""
    from quax_amd.methods.hartree_fock import restricted_hartree_fock, rhf_gradient
    mol = _c2()
    bs = BasisSet.build(mol, "even-tempered", pure=True)
    opts = {"spectral_shift": False, "conv": 1e-10, "maxit": 200}
    with _engine(mol, bs) as eng:
        E, aux = restricted_hartree_fock(eng, mol.nelectron(), dict(opts), return_aux_data=True)
        g = rhf_gradient(eng, aux)
        Em, auxm = restricted_hartree_fock(eng, mol.nelectron(),
                                           dict(opts, precision="mixed", fp32_threshold=1e-8),
                                           return_aux_data=True)
        gm = rhf_gradient(eng, auxm, precision="mixed", fp32_threshold=1e-8)
    assert abs(Em - E) < 1e-8
    assert np.abs(np.asarray(gm) - np.asarray(g)).max() < 1e-7


def test_mixed_threshold_is_validated(gpu):
    from quax_amd._lib import QuaxNativeError
    mol, bs = small_d_cluster()
    with _engine(mol, bs) as eng:
        D = _density(eng.nbf)
        for bad in (-1.0, float("nan"), float("inf")):
            with pytest.raises(QuaxNativeError):
                eng.jk(D, precision=MIXED, fp32_threshold=bad)
            with pytest.raises(QuaxNativeError):
                eng.eri_grad(D, precision=MIXED, fp32_threshold=bad)
        with pytest.raises(QuaxNativeError):
            eng.set_streams(0)
