"""RHF energy and analytic gradient on the GPU path vs the oracle and the
golden water fixture (1e-10 Eh energy, 1e-9 Eh/bohr gradient)."""
import os

import numpy as np
import pytest

import oracle
from quax_amd import workloads
from quax_amd.basis import BasisSet

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_water_energy_and_gradient_match_golden(gpu):
    from quax_amd import core
    g = np.load(os.path.join(HERE, "golden", "water_sto3g.npz"))
    mol = workloads.water_sto3g()
    opts = {"spectral_shift": False, "conv": 1e-12}
    E = core.energy(mol, "sto-3g", "hf", options=opts)
    assert abs(E - float(g["energy"])) < 1e-10
    grad = core.geom_deriv(mol, "sto-3g", "hf", options=opts)
    assert np.abs(grad - g["gradient"]).max() < 1e-9


def test_water_gradient_pieces_vs_oracle(gpu):
    """Each term of the analytic gradient at the oracle's converged density."""
    import torch
    from quax_amd.engine import EriEngine
    mol = workloads.water_sto3g()
    bs = BasisSet.build(mol, "sto-3g")
    ob = oracle.OracleBasis(mol, bs)
    S, T, V = ob.oei()
    G = ob.compute_2e_int(1e-14)
    Z = np.array([mol.charge(i) for i in range(3)])
    scf = oracle.rhf(S, T, V, G, 10, oracle.nuclear_repulsion(mol.geometry(), Z), conv=1e-13, maxit=200)
    D, C, eps = scf["D"], scf["C"], scf["eps"]
    W = C[:, :5] @ np.diag(eps[:5]) @ C[:, :5].T
    dS, dT, dV = ob.oei_deriv1()
    dG = ob.eri_deriv_core1(1e-14)
    g2_ref = 2 * np.einsum("xpqrs,pq,rs->x", dG, D, D) - np.einsum("xpqrs,pr,qs->x", dG, D, D)
    g1_ref = np.einsum("xpq,pq->x", dT + dV, 2 * D) - np.einsum("xpq,pq->x", dS, 2 * W)
    with EriEngine(mol, bs, ints_tol=1e-14) as eng:
        g2 = eng.eri_grad(D).cpu().numpy()
        g1 = eng.oei_grad(2 * D, -2 * W).cpu().numpy()
    assert np.abs(g2 - g2_ref).max() < 1e-10, (g2 - g2_ref)
    assert np.abs(g1 - g1_ref).max() < 1e-10, (g1 - g1_ref)


def test_libint_interface_mirror(gpu, tmp_path):
    from quax_amd.integrals import libint_interface as li
    from quax_amd.integrals.tei import TEI
    g = np.load(os.path.join(HERE, "golden", "water_sto3g.npz"))
    mol = workloads.water_sto3g()
    xyz = str(tmp_path / "geom.xyz")
    mol.save_xyz_file(xyz, True)
    li.initialize(xyz, "sto-3g", "sto-3g", "sto-3g", "sto-3g", 1e-14)
    try:
        G = li.compute_2e_int("eri", 0.0)
        assert G.shape == (7 ** 4,)
        assert np.abs(G.reshape(7, 7, 7, 7) - g["eri"]).max() < 1e-12
        bs = BasisSet.build(mol, "sto-3g")
        tei = TEI(bs, bs, bs, bs, xyz, 1, {}, "core")
        dv = np.zeros(9, dtype=int)
        dv[5] = 1
        d = np.asarray(tei.eri_deriv(None, dv))
        assert np.abs(d - g["eri_deriv1"][5]).max() < 1e-12
        primal, tangent = tei.eri_jvp((None,), (dv,))
        assert np.abs(np.asarray(tangent) - g["eri_deriv1"][5]).max() < 1e-12
        batch, axis = tei.eri_deriv_batch((None, np.eye(9, dtype=int)), (None, 0))
        assert axis == 0 and np.abs(np.asarray(batch) - g["eri_deriv1"]).max() < 1e-12
        d2 = li.compute_2e_deriv("eri", 0.0, dv).reshape(7, 7, 7, 7)
        assert np.abs(d2 - g["eri_deriv1"][5]).max() < 1e-12
        with pytest.raises(ValueError):
            li.compute_2e_deriv("eri", 0.0, dv[:5])
        with pytest.raises(ValueError):
            li.compute_2e_int("coulomb-ish", 0.0)
    finally:
        li.finalize()


def test_cluster_energy_vs_oracle(gpu):
    """Small even-tempered cluster with d shells: GPU RHF vs oracle RHF."""
    from quax_amd.engine import EriEngine
    from quax_amd.methods.hartree_fock import restricted_hartree_fock, rhf_gradient
    mol = workloads.cluster10_subset4()
    mol = type(mol)([mol.symbol(i) for i in range(2)], mol.geometry()[:2])   # C2, 12 electrons
    bs = BasisSet.build(mol, "even-tempered")
    ob = oracle.OracleBasis(mol, bs)
    S, T, V = ob.oei()
    G = ob.compute_2e_int(1e-14)
    Z = np.array([mol.charge(i) for i in range(mol.natom())])
    enuc = oracle.nuclear_repulsion(mol.geometry(), Z)
    ref = oracle.rhf(S, T, V, G, mol.nelectron(), enuc, conv=1e-12, maxit=300, damping=True)
    with EriEngine(mol, bs, ints_tol=1e-14) as eng:
        E, aux = restricted_hartree_fock(eng, mol.nelectron(), {"spectral_shift": False, "conv": 1e-12,
                                                                "damping": True, "maxit": 300},
                                         return_aux_data=True)
        grad = rhf_gradient(eng, aux)
    assert abs(E - ref["E"]) < 1e-10
    dS, dT, dV = ob.oei_deriv1()
    gref = oracle.rhf_gradient(ref, dS, dT, dV, ob.eri_deriv_core1(1e-14), mol.geometry(), Z)
    assert np.abs(grad - gref).max() < 1e-9
