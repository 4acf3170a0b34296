"""AO -> MO transformation (FP64 MFMA GEMM chain) against numpy on the CPU,
with the reference's index convention (energy_utils.py:46-73):
out[I,J,K,L] = sum_pqrs Ci[p,I] Cj[q,J] Ck[r,K] Cl[s,L] G[p,q,r,s]."""
import numpy as np
import pytest

from quax_amd import workloads
from quax_amd.basis import BasisSet

pytestmark = pytest.mark.gpu


def _ref(G, Ci, Cj, Ck, Cl):
    # the reference's own sequence: contract l, k, j, i, new index first
    def step(C, X):
        return np.tensordot(C, X, axes=[(0,), (3,)])
    return step(Ci, step(Cj, step(Ck, step(Cl, G))))


@pytest.mark.parametrize("n,m", [((7, 7, 7, 7), (7, 7, 7, 7)), ((5, 9, 13, 17), (3, 11, 2, 19)),
                                 ((1, 1, 1, 1), (1, 1, 1, 1)), ((33, 20, 33, 20), (5, 15, 5, 15)),
                                 ((64, 64, 64, 64), (64, 64, 64, 64)), ((70, 3, 66, 65), (81, 2, 17, 64))])
def test_partial_transformation_matches_numpy(gpu, n, m):
    from quax_amd.methods.energy_utils import partial_tei_transformation
    rng = np.random.default_rng(sum(n) + 7 * sum(m))
    G = rng.standard_normal(n)
    Cs = [rng.standard_normal((n[k], m[k])) for k in range(4)]
    out = partial_tei_transformation(G, *Cs).cpu().numpy()
    ref = _ref(G, *Cs)
    assert out.shape == tuple(m)
    assert np.abs(out - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())
    # einsum form of the same definition (energy_utils.py:57-60)
    if max(n) <= 33:
        ein = np.einsum("pqrs,pP,qQ,rR,sS->PQRS", G, *Cs, optimize=True)
        assert np.abs(out - ein).max() < 1e-12 * max(1.0, np.abs(ein).max())


def test_water_mo_integrals(gpu):
    """MO integrals of the RHF/STO-3G water: (ia|jb) for MP2 and the full
    tensor for CCSD, from the GPU ERI tensor and the GPU SCF orbitals."""
    import torch
    from quax_amd.engine import EriEngine
    from quax_amd.methods.energy_utils import partial_tei_transformation, tei_transformation
    from quax_amd.methods.hartree_fock import restricted_hartree_fock
    mol = workloads.water_sto3g()
    bs = BasisSet.build(mol, "sto-3g")
    with EriEngine(mol, bs, ints_tol=1e-14) as eng:
        G = eng.compute_2e_int()
        E, aux = restricted_hartree_fock(eng, mol.nelectron(), {"spectral_shift": False, "conv": 1e-12},
                                         return_aux_data=True)
    C = aux["C"]
    nocc = mol.nelectron() // 2
    Gn, Cn = G.cpu().numpy(), C.cpu().numpy()
    full = tei_transformation(G, C).cpu().numpy()
    assert np.abs(full - _ref(Gn, Cn, Cn, Cn, Cn)).max() < 1e-12
    ovov = partial_tei_transformation(G, C[:, :nocc], C[:, nocc:], C[:, :nocc], C[:, nocc:]).cpu().numpy()
    assert np.abs(ovov - full[:nocc, nocc:, :nocc, nocc:]).max() < 1e-12
    # MO-basis identities: 8-fold symmetry survives a real orthogonal-like transform
    assert np.abs(full - full.transpose(1, 0, 2, 3)).max() < 1e-12
    assert np.abs(full - full.transpose(2, 3, 0, 1)).max() < 1e-12
    # sum_i (ii|jj)-type check: 2 sum_ij (ii|jj) - (ij|ij) = E_2e of the closed shell
    ii = np.arange(nocc)
    e2 = 2 * full[ii[:, None], ii[:, None], ii[None, :], ii[None, :]].sum() - \
        full[ii[:, None], ii[None, :], ii[:, None], ii[None, :]].sum()
    D = aux["D"].cpu().numpy()
    e2_ao = 2 * np.einsum("pqrs,pq,rs->", Gn, D, D) - np.einsum("prqs,pq,rs->", Gn, D, D)
    assert abs(e2 - e2_ao) < 1e-11


def test_rejects_bad_shapes(gpu):
    from quax_amd.methods.energy_utils import partial_tei_transformation
    G = np.zeros((3, 3, 3, 3))
    with pytest.raises(ValueError):
        partial_tei_transformation(G, np.zeros((2, 2)), np.zeros((3, 2)), np.zeros((3, 2)), np.zeros((3, 2)))
