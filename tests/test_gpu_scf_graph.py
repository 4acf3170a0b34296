"""RHF SCF replayed from one captured HIP graph (qx_scf_*, GraphSCF).

* Parity: the converged energy against the oracle's RHF (oracle.rhf, a
  restatement of hartree_fock.py:8-119) and the golden water fixture, 1e-10
  Eh; the density against the eager driver's (host eigh) to 1e-8.
* Graph replays against the same kernels launched directly: iteration by
  iteration, replays interleaved with eager work and synchronisations, every
  replay's energy, error and density checked (graphs whose nodes misbehave
  on replays after the first are caught here).
* Batches: molecules of uneven size in one fragmented context and one graph,
  energies and gradients against single-molecule runs.
"""
import os

import numpy as np
import pytest

import oracle
from quax_amd import workloads
from quax_amd.basis import BasisSet
from quax_amd.molecule import Molecule

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _c2():
    mol = workloads.cluster10_subset4()
    return Molecule([mol.symbol(i) for i in range(2)], mol.geometry()[:2])


def _rhf(mol, bs, opts, tol=1e-14):
    from quax_amd.engine import EriEngine
    from quax_amd.methods.hartree_fock import restricted_hartree_fock
    with EriEngine(mol, bs, ints_tol=tol) as eng:
        return restricted_hartree_fock(eng, mol.nelectron(), dict(opts), return_aux_data=True)


@pytest.mark.parametrize("shift", [False, True])
def test_water_graph_scf_matches_golden_and_eager(gpu, shift):
    g = np.load(os.path.join(HERE, "golden", "water_sto3g.npz"))
    mol = workloads.water_sto3g()
    bs = BasisSet.build(mol, "sto-3g")
    conv = 1e-9 if shift else 1e-12
    base = {"spectral_shift": shift, "diis": True, "conv": conv, "maxit": 200}
    E, aux = _rhf(mol, bs, dict(base, graph=True))
    Ee, auxe = _rhf(mol, bs, base)
    print(f"  water graph SCF: E {E:.12f} in {aux['iterations']} iterations (eager {auxe['iterations']})")
    if not shift:
        assert abs(E - float(g["energy"])) < 1e-10
    assert abs(E - Ee) < 1e-10
    assert (aux["D"] - auxe["D"]).abs().max().item() < 1e-8
    assert abs(aux["iterations"] - auxe["iterations"]) <= 2


def test_c2_graph_scf_vs_oracle_and_gradient(gpu):
    """Even-tempered C2 (d shells), DIIS from the zero guess.  The converged
    graph-SCF density is checked with the oracle's integrals: idempotent in
    the S metric with tr(D S) = nocc, stationary (F D S = S D F for F built
    from the oracle's tensor), aufbau (D projects onto the nocc lowest
    eigenvectors of that F), and its energy sum (F + H) D + E_nuc equals the
    returned one.  (The reference's own iteration -- no DIIS, damping -- ends
    in a different, 6.5 Eh higher stationary state of this molecule, which
    test_gpu_rhf pins; a numpy DIIS SCF on the oracle's tensor reaches this
    one, -72.2389497380 Eh.)  Energy, density and analytic gradient also
    agree with the eager driver's."""
    from quax_amd.engine import EriEngine
    from quax_amd.methods.hartree_fock import restricted_hartree_fock, rhf_gradient
    mol = _c2()
    bs = BasisSet.build(mol, "even-tempered")
    ob = oracle.OracleBasis(mol, bs)
    S, T, V = ob.oei()
    G = ob.compute_2e_int(1e-14)
    Z = np.array([mol.charge(i) for i in range(mol.natom())])
    enuc = oracle.nuclear_repulsion(mol.geometry(), Z)
    opts = {"spectral_shift": False, "diis": True, "conv": 1e-12, "maxit": 300}
    with EriEngine(mol, bs, ints_tol=1e-14) as eng:
        E, aux = restricted_hartree_fock(eng, mol.nelectron(), dict(opts, graph=True), return_aux_data=True)
        g = rhf_gradient(eng, aux)
        Ee, auxe = restricted_hartree_fock(eng, mol.nelectron(), dict(opts), return_aux_data=True)
        ge = rhf_gradient(eng, auxe)
    nocc = mol.nelectron() // 2
    D = aux["D"].cpu().numpy()
    H = T + V
    F = H + 2 * np.einsum("pqrs,rs->pq", G, D) - np.einsum("prqs,rs->pq", G, D)
    assert np.abs(D @ S @ D - D).max() < 1e-9
    assert abs(np.trace(D @ S) - nocc) < 1e-9
    assert np.abs(F @ D @ S - S @ D @ F).max() < 1e-8
    A = np.linalg.inv(np.linalg.cholesky(S)).T
    w, C2 = np.linalg.eigh(A.T @ F @ A)
    C = A @ C2
    assert np.abs(C[:, :nocc] @ C[:, :nocc].T - D).max() < 1e-8
    assert abs(np.sum((F + H) * D) + enuc - E) < 1e-9
    assert abs(E - -72.2389497380) < 1e-8
    assert abs(E - Ee) < 1e-10
    assert (aux["D"] - auxe["D"]).abs().max().item() < 1e-8
    assert np.abs(g - ge).max() < 1e-8


def test_ill_conditioned_overlap_finishes_eagerly(gpu):
    """Water in the even-tempered basis (lambda_min(S) = 4.5e-5, |A|^2 ~ 2e4):
    the purified density's rounding floor (DIIS error ~1e-10) is above
    conv = 1e-11, so the graph iterations stall and eigh-based iterations
    from the graph's density finish; the result equals the eager driver's."""
    mol = workloads.water_sto3g()
    bs = BasisSet.build(mol, "even-tempered", pure=True)
    base = {"spectral_shift": False, "diis": True, "conv": 1e-11, "maxit": 200}
    E, aux = _rhf(mol, bs, dict(base, graph=True), tol=1e-12)
    Ee, auxe = _rhf(mol, bs, base, tol=1e-12)
    print(f"  water even-tempered: graph {aux['graph_iterations']} + eager "
          f"{aux['iterations'] - aux['graph_iterations']} iterations, eager alone {auxe['iterations']}")
    assert 0 < aux["graph_iterations"] < aux["iterations"]
    assert aux["drms"] <= 1e-11
    assert abs(E - Ee) < 1e-10
    assert (aux["D"] - auxe["D"]).abs().max().item() < 1e-8


def _graph_state(mol, bs, opts):
    import torch
    from quax_amd.engine import EriEngine
    from quax_amd.methods.hartree_fock import nuclear_repulsion
    from quax_amd.methods.scf_graph import GraphSCF
    eng = EriEngine(mol, bs, ints_tol=1e-12)
    S, T, V = eng.oei()
    H = T + V
    A = torch.linalg.inv(torch.linalg.cholesky(S)).T.contiguous()
    Z = [mol.charge(i) for i in range(mol.natom())]
    scf = GraphSCF(eng, [H], [S], [A], [mol.nelectron() // 2], [nuclear_repulsion(mol.geometry(), Z)], **opts)
    scf.init([torch.zeros_like(H)])
    return eng, scf


def test_graph_replays_match_direct_launches(gpu):
    """Ten iterations of a 3-atom even-tempered cluster (DIIS on): one SCF
    state stepped by direct launches, another by graph replays with eager
    work and synchronisations between them; each iteration's energy, DIIS
    error and density must agree (only the order of the J/K kernels' FP64
    atomic additions differs)."""
    import torch
    from quax_amd.methods.scf_graph import SCF_D
    m = workloads.cluster10()
    mol = Molecule([m.symbol(i) for i in range(3)], m.geometry()[:3])
    bs = BasisSet.build(mol, "even-tempered", pure=True)
    opts = dict(diis=True, diis_space=8, conv=0.0, maxit=50, purify_iters=100)
    e1, direct = _graph_state(mol, bs, opts)
    e2, graphed = _graph_state(mol, bs, opts)
    try:
        for it in range(10):
            direct.step(1, use_graph=False)
            graphed.step(1, use_graph=True)
            # eager work on the same device between replays
            x = torch.randn(512, 512, dtype=torch.float64, device="cuda")
            _ = (x @ x).sum().item()
            a, b = direct.status(), graphed.status()
            Da, Db = direct.matrix(SCF_D)[0], graphed.matrix(SCF_D)[0]
            dD = (Da - Db).abs().max().item()
            print(f"  iteration {it + 1}: E {a['E'][0]:.12f} / {b['E'][0]:.12f}, drms {a['drms'][0]:.2e}, "
                  f"|dD| {dD:.1e}", flush=True)
            assert a["iterations"][0] == b["iterations"][0] == it + 1
            assert abs(a["E"][0] - b["E"][0]) < 1e-9
            assert abs(a["drms"][0] - b["drms"][0]) < 1e-9 + 1e-6 * a["drms"][0]
            assert dD < 1e-7
            assert abs(a["idempotency"][0]) < 1e-8 and abs(b["idempotency"][0]) < 1e-8
    finally:
        direct.close()
        graphed.close()
        e1.close()
        e2.close()


def test_batched_graph_scf_matches_single_molecules(gpu):
    """Water, C2 and H2 (even-tempered, spherical; 84, 56 and 56
    functions) as fragments of one context, all iterated by one graph;
    energies and gradients against each molecule run alone through the eager
    driver."""
    from quax_amd.engine import EriEngine
    from quax_amd.methods.batch import batched_rhf
    from quax_amd.methods.hartree_fock import restricted_hartree_fock, rhf_gradient
    mols = [workloads.water_sto3g(), _c2(), Molecule(["H", "H"], [[0.0, 0.0, 0.0], [0.0, 0.0, 1.4]])]
    E, grads, aux = batched_rhf(mols, "even-tempered", pure=True, ints_tol=1e-12, conv=1e-11, maxit=200,
                                graph=True)
    print(f"  batched graph SCF: iterations {aux['iterations']}, converged {aux['converged']}")
    assert all(aux["converged"])
    for k, m in enumerate(mols):
        bs = BasisSet.build(m, "even-tempered", pure=True)
        opts = {"spectral_shift": False, "diis": True, "conv": 1e-11, "maxit": 200}
        with EriEngine(m, bs, ints_tol=1e-12) as eng:
            Ek, auxk = restricted_hartree_fock(eng, m.nelectron(), opts, return_aux_data=True)
            gk = rhf_gradient(eng, auxk)
        assert abs(E[k] - Ek) < 1e-9, (k, E[k], Ek)
        assert np.abs(grads[k] - gk).max() < 1e-7


def test_graph_scf_rejects_bad_arguments(gpu):
    from quax_amd._lib import QuaxNativeError
    from quax_amd.engine import EriEngine
    from quax_amd.methods.hartree_fock import restricted_hartree_fock
    from quax_amd.methods.scf_graph import GraphSCF
    mol = workloads.water_sto3g()
    bs = BasisSet.build(mol, "sto-3g")
    with EriEngine(mol, bs) as eng:
        S, T, V = eng.oei()
        H = T + V
        with pytest.raises(QuaxNativeError):
            GraphSCF(eng, [H], [S], [S], [50], [0.0])            # more occupied orbitals than functions
        with pytest.raises(QuaxNativeError):
            GraphSCF(eng, [H], [S], [S], [5], [0.0], diis_space=64)
        with pytest.raises(QuaxNativeError):
            GraphSCF(eng, [H], [S], [S], [5], [0.0], offsets=[3])  # block outside the context
        with pytest.raises(ValueError):
            restricted_hartree_fock(eng, 10, {"graph": True, "damping": True})
