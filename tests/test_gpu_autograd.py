"""The adjoint kernel behind torch.autograd: E2(R, D) = sum D (2J - K)
differentiated by backward() against central finite differences of the
same energy (step 1e-4 bohr in R, 1e-5 in D)."""
import numpy as np
import pytest

from quax_amd import workloads
from quax_amd.basis import BasisSet
from quax_amd.molecule import Molecule

pytestmark = pytest.mark.gpu


def _energy_at(mol, bs, xyz, D):
    """E2 at displaced coordinates: shells reference atoms, so the same basis
    follows the moved nuclei."""
    import torch
    from quax_amd.engine import EriEngine
    m = Molecule([mol.symbol(i) for i in range(mol.natom())], xyz)
    with EriEngine(m, bs, ints_tol=1e-14) as eng:
        J, K = eng.jk(D)
        return float(torch.sum(torch.as_tensor(D, device=J.device) * (2 * J - K)))


def _density(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n)) * 0.3
    return X @ X.T / n


@pytest.mark.parametrize("case", ["water", "c2"])
def test_backward_matches_finite_differences(gpu, case):
    import torch
    from quax_amd.autograd import eri_energy
    from quax_amd.engine import EriEngine
    if case == "water":
        mol, bname = workloads.water_sto3g(), "sto-3g"
    else:
        m = workloads.cluster10()
        mol, bname = Molecule([m.symbol(i) for i in range(2)], m.geometry()[:2]), "even-tempered"
    bs = BasisSet.build(mol, bname)
    D = _density(bs.nbf(), 3)
    xyz0 = np.asarray(mol.geometry(), dtype=np.float64).reshape(-1, 3)
    with EriEngine(mol, bs, ints_tol=1e-14) as eng:
        geom = torch.tensor(xyz0, dtype=torch.float64, requires_grad=True)
        Dt = torch.tensor(D, dtype=torch.float64, device="cuda", requires_grad=True)
        E = eri_energy(eng, geom, Dt)
        E.backward()
        g = geom.grad.numpy().reshape(-1)
        gD = Dt.grad.cpu().numpy()
        J, K = eng.jk(D)
        E0 = float(torch.sum(torch.as_tensor(D, device="cuda") * (2 * J - K)))
    assert abs(float(E.detach()) - E0) < 1e-12 * max(1.0, abs(E0))
    # geometry: central differences, every coordinate (basis built per geometry)
    h = 1e-4
    fd = np.zeros_like(g)
    for k in range(g.size):
        xp, xm = xyz0.copy().reshape(-1), xyz0.copy().reshape(-1)
        xp[k] += h
        xm[k] -= h
        Ep = _energy_at(mol, bs, xp.reshape(-1, 3), D)
        Em = _energy_at(mol, bs, xm.reshape(-1, 3), D)
        fd[k] = (Ep - Em) / (2 * h)
    assert np.abs(g - fd).max() < 2e-7 * max(1.0, np.abs(fd).max())
    assert np.abs(g.reshape(-1, 3).sum(0)).max() < 1e-10          # translational invariance
    # density: directional derivative along a symmetric direction
    rng = np.random.default_rng(9)
    X = rng.standard_normal(D.shape)
    X = X + X.T
    eps = 1e-5
    with EriEngine(mol, bs, ints_tol=1e-14) as eng:
        def e2(Dm):
            J, K = eng.jk(Dm)
            return float(torch.sum(torch.as_tensor(Dm, device="cuda") * (2 * J - K)))
        fdD = (e2(D + eps * X) - e2(D - eps * X)) / (2 * eps)
    assert abs(float(np.sum(gD * X)) - fdD) < 1e-7 * max(1.0, abs(fdD))


def test_geometry_mismatch_is_rejected(gpu):
    import torch
    from quax_amd.autograd import eri_energy
    from quax_amd.engine import EriEngine
    mol = workloads.water_sto3g()
    bs = BasisSet.build(mol, "sto-3g")
    with EriEngine(mol, bs) as eng:
        geom = torch.tensor(np.asarray(mol.geometry()).reshape(-1, 3) + 0.1, dtype=torch.float64, requires_grad=True)
        with pytest.raises(ValueError):
            eri_energy(eng, geom, torch.eye(bs.nbf(), dtype=torch.float64, device="cuda"))
