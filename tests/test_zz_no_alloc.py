"""qx_jk_build_noalloc (J/K with no device allocation: work matrices in the
library's file-scope scratch array) against qx_jk_build, called through the
C ABI.  Two calls in a row on one GPU, the second on a smaller molecule, so a
stale or uncleared scratch array would show; each result within 1e-12
(relative to the largest element) of the allocating entry point, whose
parity with the oracle the J/K tests establish."""
import ctypes

import numpy as np
import pytest

from quax_amd import workloads
from quax_amd.basis import BasisSet
from quax_amd.molecule import Molecule

pytestmark = pytest.mark.gpu


def _density(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n)) * 0.3
    return X @ X.T / n


def _noalloc_jk(torch, eng, D):
    from quax_amd import _lib
    Dt = torch.as_tensor(D, dtype=torch.float64, device=eng.device).contiguous()
    J = torch.full((eng.nbf, eng.nbf), float("nan"), dtype=torch.float64, device=eng.device)
    K = torch.full((eng.nbf, eng.nbf), float("nan"), dtype=torch.float64, device=eng.device)
    stream = ctypes.c_void_p(torch.cuda.current_stream(eng.device).cuda_stream)
    rc = eng.L.qx_jk_build_noalloc(eng.ctx, _lib.ptr(Dt), _lib.ptr(J), _lib.ptr(K), stream)
    return rc, J.cpu().numpy(), K.cpu().numpy()


def _cases():
    # larger first, then smaller: spherical subset4 cluster (112 AO, 120
    # Cartesian), then Cartesian water/STO-3G (7)
    sub = workloads.cluster10_subset4()
    water = workloads.water_sto3g()
    return [(sub, BasisSet.build(sub, "even-tempered", pure=True), 1e-12),
            (water, BasisSet.build(water, "sto-3g"), 1e-14)]


def test_noalloc_jk_matches_jk_build_twice(gpu):
    torch = gpu
    from quax_amd.engine import EriEngine
    for k, (mol, bs, tol) in enumerate(_cases()):
        print(f"  no-alloc J/K call {k + 1}: {mol.natom()} atoms, {bs.nbf()} functions", flush=True)
        with EriEngine(mol, bs, ints_tol=tol) as eng:
            D = _density(eng.nbf, 5 + k)
            rc, Jn, Kn = _noalloc_jk(torch, eng, D)
            assert rc == 0, eng.L.qx_last_error(eng.ctx).decode()
            J, K = (x.cpu().numpy() for x in eng.jk(D))
        sj = max(1.0, np.abs(J).max())
        sk = max(1.0, np.abs(K).max())
        assert np.abs(Jn - J).max() <= 1e-12 * sj, np.abs(Jn - J).max()
        assert np.abs(Kn - K).max() <= 1e-12 * sk, np.abs(Kn - K).max()


def test_noalloc_rejects_contexts_beyond_the_scratch_array(gpu):
    """A Cartesian context with more than 1254 functions needs more than the
    3 Mi-double array: the call fails with QX_EINVAL and a message, and
    launches nothing."""
    torch = gpu
    from quax_amd import _lib
    from quax_amd.engine import EriEngine
    cap = int(_lib.load().qx_jk_noalloc_capacity())
    assert cap == 3 << 20
    rng = np.random.default_rng(3)
    natom = 43                                     # 43 x 30 = 1290 Cartesian functions
    xyz = np.stack(np.meshgrid(np.arange(4), np.arange(4), np.arange(3), indexing="ij"), -1).reshape(-1, 3)[:natom]
    mol = Molecule(["C"] * natom, 3.5 * xyz + rng.uniform(-0.1, 0.1, (natom, 3)))
    bs = BasisSet.build(mol, "even-tempered")
    with EriEngine(mol, bs, ints_tol=1e-10) as eng:
        assert 2 * eng.ncart ** 2 > cap
        rc, _, _ = _noalloc_jk(torch, eng, np.eye(eng.nbf))
        assert rc == _lib.QX_EINVAL
        assert "scratch array" in eng.L.qx_last_error(eng.ctx).decode()
