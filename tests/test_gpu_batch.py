"""Batched form: molecules as fragments of one context (one launch per class
serves them all) against one context per molecule, and the lock-step
batched RHF against restricted_hartree_fock + rhf_gradient per molecule."""
import numpy as np
import pytest

from quax_amd import workloads
from quax_amd.basis import BasisSet, Shell
from quax_amd.molecule import Molecule

pytestmark = pytest.mark.gpu


def _smallest(k):
    ms = workloads.batch256()
    return sorted(ms, key=lambda m: (m.natom(), m.nelectron()))[:k]


def _combined(mols, bases):
    symbols, coords, shells, frag = [], [], [], []
    aoff = 0
    for k, (m, b) in enumerate(zip(mols, bases)):
        symbols += [m.symbol(i) for i in range(m.natom())]
        coords.append(np.asarray(m.geometry()).reshape(-1, 3))
        shells += [Shell(s.l, s.atom + aoff, list(s.exps), list(s.coefs)) for s in b.shells]
        frag += [k] * m.natom()
        aoff += m.natom()
    return Molecule(symbols, np.vstack(coords)), BasisSet(bases[0].basis_name, shells, pure=bases[0].pure), frag


def test_fragment_context_equals_separate_contexts(gpu):
    import torch
    from quax_amd.engine import EriEngine
    mols = _smallest(3)
    bases = [BasisSet.build(m, "even-tempered", pure=True) for m in mols]
    comb, cb, frag = _combined(mols, bases)
    nb = [b.nbf() for b in bases]
    off = np.concatenate([[0], np.cumsum(nb)]).astype(int)
    rng = np.random.default_rng(4)
    Ds = []
    for n in nb:
        X = rng.standard_normal((n, n)) * 0.3
        Ds.append(X @ X.T / n)
    D = np.zeros((off[-1], off[-1]))
    for k, Dk in enumerate(Ds):
        D[off[k]:off[k + 1], off[k]:off[k + 1]] = Dk
    with EriEngine(comb, cb, ints_tol=1e-12, fragments=frag) as eng:
        J, K = (x.cpu().numpy() for x in eng.jk(D))
        g = eng.eri_grad(D).cpu().numpy()
        nq = eng.num_quartets()
    nat = np.concatenate([[0], np.cumsum([m.natom() for m in mols])]).astype(int)
    total = 0
    for k, (m, b) in enumerate(zip(mols, bases)):
        with EriEngine(m, b, ints_tol=1e-12) as e1:
            Jk, Kk = (x.cpu().numpy() for x in e1.jk(Ds[k]))
            gk = e1.eri_grad(Ds[k]).cpu().numpy()
            total += e1.num_quartets()
        s = slice(off[k], off[k + 1])
        assert np.abs(J[s, s] - Jk).max() < 1e-11
        assert np.abs(K[s, s] - Kk).max() < 1e-11
        assert np.abs(g[3 * nat[k]:3 * nat[k + 1]] - gk).max() < 1e-10
        # nothing leaks between molecules
        other = np.ones(off[-1], dtype=bool)
        other[s] = False
        assert np.abs(J[s][:, other]).max() == 0.0 and np.abs(K[s][:, other]).max() == 0.0
    assert nq == total


def _well_behaved():
    """Closed-shell molecules with one clear SCF ground state (random
    synthetic geometries can have several SCF solutions, which DIIS may
    reach along different paths)."""
    return [Molecule(["O", "H", "H"], [[0.0, 0.0, 0.0], [0.0, 1.43, 1.1], [0.0, -1.43, 1.1]]),
            Molecule(["H", "H"], [[0.0, 0.0, 0.0], [0.0, 0.0, 1.4]]),
            Molecule(["C", "O"], [[0.0, 0.0, 0.0], [0.0, 0.0, 2.13]])]


def test_batched_rhf_matches_single_molecule_runs(gpu):
    from quax_amd.engine import EriEngine
    from quax_amd.methods.batch import batched_rhf
    from quax_amd.methods.hartree_fock import restricted_hartree_fock, rhf_gradient
    mols = _well_behaved()
    E, G, aux = batched_rhf(mols, "even-tempered", pure=True, ints_tol=1e-12, conv=1e-10)
    assert all(aux["converged"])
    for k, m in enumerate(mols):
        b = BasisSet.build(m, "even-tempered", pure=True)
        with EriEngine(m, b, ints_tol=1e-12) as e1:
            Ek, auxk = restricted_hartree_fock(e1, m.nelectron(), {"diis": True, "spectral_shift": False,
                                                                  "conv": 1e-10, "maxit": 200},
                                               return_aux_data=True)
            gk = rhf_gradient(e1, auxk)
        assert abs(E[k] - Ek) < 1e-9
        assert np.abs(G[k] - gk).max() < 1e-7


def test_masked_jk_leaves_out_fragments(gpu):
    """qx_jk_build_masked: the blocks of masked fragments come out zero, the
    others equal the unmasked pass."""
    from quax_amd._lib import QuaxNativeError
    from quax_amd.engine import EriEngine
    mols = _smallest(3)
    bases = [BasisSet.build(m, "even-tempered", pure=True) for m in mols]
    comb, cb, frag = _combined(mols, bases)
    nb = [b.nbf() for b in bases]
    off = np.concatenate([[0], np.cumsum(nb)]).astype(int)
    rng = np.random.default_rng(8)
    D = np.zeros((off[-1], off[-1]))
    for k, n in enumerate(nb):
        X = rng.standard_normal((n, n)) * 0.3
        D[off[k]:off[k + 1], off[k]:off[k + 1]] = X @ X.T / n
    with EriEngine(comb, cb, ints_tol=1e-12, fragments=frag) as eng:
        J, K = (x.cpu().numpy() for x in eng.jk(D))
        for mask in ([0, 1, 0], [1, 0, 1], [1, 1, 1], [0, 0, 0]):
            Jm, Km = (x.cpu().numpy() for x in eng.jk(D, frag_skip=mask))
            for k in range(3):
                s = slice(off[k], off[k + 1])
                if mask[k]:
                    assert np.abs(Jm[s, s]).max() == 0.0 and np.abs(Km[s, s]).max() == 0.0
                else:
                    assert np.abs(Jm[s, s] - J[s, s]).max() < 1e-12
                    assert np.abs(Km[s, s] - K[s, s]).max() < 1e-12
    # a context without fragments takes no mask
    m = mols[0]
    with EriEngine(m, bases[0], ints_tol=1e-12) as e1:
        with pytest.raises(QuaxNativeError):
            e1.jk(D[:nb[0], :nb[0]], frag_skip=[0])


def test_class_tiles_do_not_change_results(gpu):
    """Every quartet tile (threads per workgroup) of the streaming kernels
    gives the same J, K and gradient; invalid tiles are refused."""
    from quax_amd._lib import QuaxNativeError
    from quax_amd.engine import EriEngine
    m = workloads.cluster10()
    mol = Molecule([m.symbol(i) for i in range(3)], m.geometry()[:3])
    bs = BasisSet.build(mol, "even-tempered", pure=True)
    with EriEngine(mol, bs, ints_tol=1e-12) as eng:
        n = eng.nbf
        X = np.random.default_rng(3).standard_normal((n, n)) * 0.3
        D = X @ X.T / n
        default = [eng.class_tiles(p) for p in (0, 1)]
        assert all(((t >= 64) & (t <= 256) & (t % 64 == 0)).all() for t in default)
        J, K = (x.cpu().numpy() for x in eng.jk(D))
        g = eng.eri_grad(D).cpu().numpy()
        nc = len(default[0])
        rng = np.random.default_rng(5)
        for tiles in ([64] * nc, [192] * nc, [256] * nc, list(rng.choice([64, 128, 192, 256], nc))):
            for p in (0, 1):
                eng.set_class_tiles(p, tiles)
            Jt, Kt = (x.cpu().numpy() for x in eng.jk(D))
            gt = eng.eri_grad(D).cpu().numpy()
            assert np.abs(Jt - J).max() < 1e-12 and np.abs(Kt - K).max() < 1e-12
            assert np.abs(gt - g).max() < 1e-12
        for p in (0, 1):
            eng.set_class_tiles(p, None)
            assert (eng.class_tiles(p) == default[p]).all()
        # the per-class wave reduction of the gradient terms changes no result
        wdef = eng.grad_wave_reduce()
        for flags in (np.ones(nc, dtype=bool), np.zeros(nc, dtype=bool)):
            eng.set_grad_wave_reduce(flags)
            assert (eng.grad_wave_reduce() == flags).all()
            assert np.abs(eng.eri_grad(D).cpu().numpy() - g).max() < 1e-12
        eng.set_grad_wave_reduce(None)
        assert (eng.grad_wave_reduce() == wdef).all()
        for bad in ([32] * nc, [320] * nc, [100] * nc):
            with pytest.raises(QuaxNativeError):
                eng.set_class_tiles(0, bad)


def test_batched_graph_scf_seeded_molecules(gpu):
    """Seeded batch molecules (3-12 atoms, C/H/O) through the graph SCF
    (converged molecules masked out of J/K) against each molecule alone
    through the eager driver: energies 1e-9 Eh, gradients 1e-7.  Only
    molecules that both drivers converge within 40 iterations are compared:
    on some synthetic geometries DIIS from the zero guess wanders for 50-200
    iterations and, depending on rounding-level differences along the way,
    ends in different stationary states (an H-C-H fragment of the set reaches
    -37.0701200 Eh in numpy and in the eager driver and a state 1.0e-3 Eh
    lower in the graph SCF; a 3-atom one converged in ~60 iterations to
    states 0.037 Eh apart on one box and to the same state on another), so
    those molecules compare no implementation."""
    from quax_amd.engine import EriEngine
    from quax_amd.methods.batch import batched_rhf
    from quax_amd.methods.hartree_fock import restricted_hartree_fock, rhf_gradient
    opts = {"diis": True, "spectral_shift": False, "conv": 1e-10, "maxit": 200}
    ref = []
    for m in _well_behaved() + _smallest(8):
        b = BasisSet.build(m, "even-tempered", pure=True)
        with EriEngine(m, b, ints_tol=1e-12) as e1:
            Ek, auxk = restricted_hartree_fock(e1, m.nelectron(), opts, return_aux_data=True)
            if auxk["drms"] > 1e-10 or abs(auxk["dE"]) > 1e-10 or auxk["iterations"] > 40:
                print(f"  {m.natom()} atoms: eager SCF took {auxk['iterations']} iterations "
                      f"(RMS error {auxk['drms']:.1e}), left out", flush=True)
                continue
            ref.append((m, Ek, rhf_gradient(e1, auxk)))
    assert len(ref) >= 6
    mols = [r[0] for r in ref]
    E, G, aux = batched_rhf(mols, "even-tempered", pure=True, ints_tol=1e-12, conv=1e-10, graph=True)
    print(f"  graph batch: iterations {aux['iterations']}, converged {aux['converged']}", flush=True)
    assert all(aux["converged"])
    compared = 0
    for k, (m, Ek, gk) in enumerate(ref):
        if aux["iterations"][k] > 40:
            print(f"  molecule {k} ({m.natom()} atoms): graph SCF took {aux['iterations'][k]} iterations, "
                  f"not compared", flush=True)
            continue
        compared += 1
        print(f"  molecule {k} ({m.natom()} atoms): dE {E[k] - Ek:.1e}, max dg {np.abs(G[k] - gk).max():.1e}",
              flush=True)
        assert abs(E[k] - Ek) < 1e-9
        assert np.abs(G[k] - gk).max() < 1e-7
    assert compared >= 6


def test_gradient_window_paths_agree(gpu):
    """The adjoint kernels accumulate a block's gradient terms in LDS over
    the atom window its quartets touch, or in global memory when the window
    is wider than the limit: a limit of 1 or 2 atoms sends all or some
    blocks down the global path, and the gradient must not change beyond
    summation order (nor the forward-over-reverse directional derivative,
    whose tangent kernels accumulate all 3N slots and ignore the limit)."""
    from test_gpu_eri import small_d_cluster
    from quax_amd.engine import EriEngine
    mol, basis = small_d_cluster()
    rng = np.random.default_rng(11)
    n = basis.nbf()
    X = rng.standard_normal((n, n)) * 0.3
    D = X @ X.T / n
    v = rng.standard_normal(3 * mol.natom())
    with EriEngine(mol, basis, ints_tol=1e-14) as eng:
        assert eng.grad_window() == 256
        g0, h0 = eng.eri_grad(D).cpu().numpy(), eng.eri_hess_dir(D, v).cpu().numpy()
        for atoms in (1, 2):
            eng.set_grad_window(atoms)
            assert eng.grad_window() == atoms
            g, h = eng.eri_grad(D).cpu().numpy(), eng.eri_hess_dir(D, v).cpu().numpy()
            print(f"  window {atoms}: max dg {np.abs(g - g0).max():.1e}, max dh {np.abs(h - h0).max():.1e}",
                  flush=True)
            assert np.abs(g - g0).max() <= 1e-12 * max(1.0, np.abs(g0).max())
            assert np.abs(h - h0).max() <= 1e-12 * max(1.0, np.abs(h0).max())
        eng.set_grad_window(0)
        assert eng.grad_window() == 256
        with pytest.raises(RuntimeError):
            eng.set_grad_window(257)


def test_fragment_gradient_beyond_256_atoms(gpu):
    """96 molecules (288 atoms, more than the 256 an LDS window holds) as
    fragments of one context: every block's window stays within one or two
    molecules, and each molecule's gradient equals its own context's."""
    from quax_amd.engine import EriEngine
    protos = [Molecule(["O", "H", "H"], [[0.0, 0.0, 0.0], [0.0, 1.43, 1.1], [0.0, -1.43, 1.1]]),
              Molecule(["O", "H", "H"], [[0.1, 0.0, 0.0], [0.0, 1.52, 0.9], [0.2, -1.31, 1.2]]),
              Molecule(["C", "O", "H"], [[0.0, 0.0, 0.0], [0.0, 0.0, 2.2], [0.0, 1.9, -0.6]]),
              Molecule(["H", "O", "H"], [[1.2, 0.3, 0.0], [0.0, 0.0, 0.0], [-0.9, 1.0, 0.3]])]
    rng = np.random.default_rng(5)
    pb = [BasisSet.build(m, "sto-3g") for m in protos]
    pD = []
    for b in pb:
        X = rng.standard_normal((b.nbf(), b.nbf())) * 0.3
        pD.append(X @ X.T / b.nbf())
    mols, which = [], []
    for k in range(96):
        p = protos[k % 4]
        shift = np.array([20.0 * (k % 6), 20.0 * ((k // 6) % 4), 20.0 * (k // 24)])
        mols.append(Molecule([p.symbol(i) for i in range(3)], np.asarray(p.geometry()).reshape(-1, 3) + shift))
        which.append(k % 4)
    bases = [BasisSet.build(m, "sto-3g") for m in mols]
    comb, cb, frag = _combined(mols, bases)
    assert comb.natom() == 288
    nb = [b.nbf() for b in bases]
    off = np.concatenate([[0], np.cumsum(nb)]).astype(int)
    D = np.zeros((off[-1], off[-1]))
    for k, w in enumerate(which):
        D[off[k]:off[k + 1], off[k]:off[k + 1]] = pD[w]
    with EriEngine(comb, cb, ints_tol=1e-12, fragments=frag) as eng:
        g = eng.eri_grad(D).cpu().numpy()
    ref = []
    for m, b, Dk in zip(protos, pb, pD):
        with EriEngine(m, b, ints_tol=1e-12) as e1:
            ref.append(e1.eri_grad(Dk).cpu().numpy())
    err = max(np.abs(g[9 * k:9 * k + 9] - ref[w]).max() for k, w in enumerate(which))
    print(f"  288 atoms in 96 fragments: max gradient difference {err:.1e}", flush=True)
    assert err < 1e-10
