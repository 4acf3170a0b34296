"""Host-side checks (no GPU): the index contract, basis data, synthetic
workloads, the spherical transform, and that the C-ABI library loads and
exports every entry point include/quax_amd.h declares."""
import ctypes
import itertools
import json
import math
import os
import re

import numpy as np
import pytest

import oracle
from quax_amd import utils, workloads
from quax_amd.basis import BasisSet, cart_overlap_shared_norm, sph_transform
from quax_amd.molecule import Molecule, read_xyz

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_how_many_derivs():
    for k in range(1, 6):
        for n in range(0, 4):
            assert utils.how_many_derivs(k, n) == oracle.how_many_derivs(k, n) == math.comb(3 * k + n - 1, n)


@pytest.mark.parametrize("natom,order", [(1, 1), (2, 1), (3, 1), (2, 2), (3, 2), (2, 3)])
def test_deriv_vec_index_matches_enumeration(natom, order):
    dim = 3 * natom
    for idx, c in enumerate(itertools.combinations_with_replacement(range(dim), order)):
        dv = np.zeros(dim, dtype=int)
        for i in c:
            dv[i] += 1
        assert utils.get_deriv_vec_idx(dv) == idx == oracle.get_deriv_vec_idx(dv)


def test_multi_index_lookup():
    assert utils.generate_multi_index_lookup(12, 1) == [[i] for i in range(12)]
    lk = utils.generate_multi_index_lookup(12, 2)
    assert len(lk) == 78 and lk[0] == [0, 0] and lk[1] == [0, 1] and lk[-1] == [11, 11]


def test_native_lookup_and_counts():
    from quax_amd import _lib
    L = _lib.load()
    for n, k in [(9, 1), (12, 2), (6, 3)]:
        cnt = L.qx_generate_multi_index_lookup(n, k, None, 0)
        buf = np.zeros(cnt * k, dtype=np.int32)
        L.qx_generate_multi_index_lookup(n, k, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), cnt)
        assert buf.reshape(cnt, k).tolist() == utils.generate_multi_index_lookup(n, k)
        assert L.qx_how_many_derivs(n // 3, k) == utils.how_many_derivs(n // 3, k)


def test_library_exports_every_declared_symbol():
    from quax_amd import _lib
    L = _lib.load()
    header = open(os.path.join(ROOT, "include", "quax_amd.h")).read()
    names = sorted(set(re.findall(r"\b(qx_[a-z0-9_]+)\s*\(", header)))
    assert len(names) >= 20
    for nm in names:
        assert hasattr(L, nm), nm
        assert nm in _lib.SIGNATURES, nm


def test_library_rejects_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from quax_amd._lib import NativeUnavailable
    from quax_amd.engine import EriEngine
    mol = workloads.water_sto3g()
    with pytest.raises(NativeUnavailable):
        EriEngine(mol, BasisSet.build(mol, "sto-3g"))


_ENOMEM_PROBE = r"""
import ctypes, resource, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from quax_amd import _lib
L = _lib.load()
vm = [int(l.split()[1]) * 1024 for l in open('/proc/self/status') if l.startswith('VmSize:')][0]
lim = vm + (1 << 30)
resource.setrlimit(resource.RLIMIT_AS, (lim, lim))
assert resource.getrlimit(resource.RLIMIT_AS)[0] == lim
dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
xyz, Z, ex, co = np.zeros(3), np.ones(1), np.ones(1), np.ones(1)
li = np.zeros(1, dtype=np.int32); at = np.zeros(1, dtype=np.int32); npr = np.ones(1, dtype=np.int32)
ctx = ctypes.c_void_p()
# 2**31 - 1 atoms: the per-atom host table alone needs 8.6 GB, far above the limit
rc = L.qx_initialize(2**31 - 1, xyz.ctypes.data_as(dp), Z.ctypes.data_as(dp), 1, li.ctypes.data_as(ip),
                     at.ctypes.data_as(ip), npr.ctypes.data_as(ip), ex.ctypes.data_as(dp), co.ctypes.data_as(dp),
                     0, 1e-12, 0, ctypes.byref(ctx))
print(rc, ctx.value, L.qx_last_error(None).decode())
"""


def test_host_allocation_failure_is_a_status_not_an_abort():
    """A C++ exception inside an entry point comes back as a status code
    (QX_ENOMEM for a failed host allocation); nothing reaches std::terminate."""
    import subprocess
    import sys
    from quax_amd import _lib
    _lib.load()
    r = subprocess.run([sys.executable, "-c", _ENOMEM_PROBE, ROOT], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    rc, ptr, msg = r.stdout.strip().split(" ", 2)
    assert int(rc) == _lib.QX_ENOMEM and ptr == "None"
    assert "allocation" in msg


def test_class_table_static_info():
    from quax_amd import _lib
    L = _lib.load()
    nc = L.qx_num_classes()
    assert nc == 21
    seen = set()
    for k in range(nc):
        l4 = (ctypes.c_int * 4)()
        npart = ctypes.c_int()
        assert L.qx_class_info(k, l4, ctypes.byref(npart)) == 0
        la, lb, lc, ld = l4[:]
        assert la >= lb and lc >= ld and (la * (la + 1) // 2 + lb) >= (lc * (lc + 1) // 2 + ld)
        assert 1 <= npart.value <= L.qx_max_parts()
        seen.add((la, lb, lc, ld))
        ef = ctypes.c_double()
        pf = (ctypes.c_double * L.qx_max_parts())()
        pc = (ctypes.c_int * L.qx_max_parts())()
        assert L.qx_class_flops(k, ctypes.byref(ef), pf, pc) == 0
        assert ef.value > 0 and all(pf[i] > 0 for i in range(npart.value))
        ng = ctypes.c_int()
        gf = (ctypes.c_double * L.qx_max_parts())()
        gp = (ctypes.c_int * L.qx_max_parts())()
        assert L.qx_class_groups(k, ctypes.byref(ng), gf, gp) == 0
        assert 1 <= ng.value <= npart.value and sum(gp[i] for i in range(ng.value)) == npart.value
        assert abs(sum(gf[i] for i in range(ng.value)) - sum(pf[i] for i in range(npart.value))) < 1e-9 * sum(gf[i] for i in range(ng.value))
        nj = ctypes.c_int()
        jf = (ctypes.c_double * 8)()
        assert L.qx_class_jk(k, ctypes.byref(nj), jf) == 0
        assert 1 <= nj.value <= 6 and all(jf[i] > ef.value / 8 for i in range(nj.value))
    assert len(seen) == 21


def test_profile_slots_are_distinct():
    from quax_amd import _lib
    L = _lib.load()
    seen = set()
    for slot in range(L.qx_profile_slots()):
        v = [ctypes.c_int() for _ in range(4)]
        assert L.qx_profile_slot(slot, *[ctypes.byref(x) for x in v]) == 0
        key = tuple(x.value for x in v)
        assert key not in seen and key[0] in (0, 1, 2) and key[3] in (0, 1) and 0 <= key[1] < 21
        seen.add(key)
    assert L.qx_profile_slot(L.qx_profile_slots(), None, None, None, None) != 0


def test_derivative_parts_partition_the_outputs():
    """Every class's derivative parts cover each (a, b, c) output triple once."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "quax_amd", "csrc"))
    import gen_eri as G
    _, classes = G.canonical_classes(2)
    for cls in classes:
        parts = G.deriv_parts(cls)
        na, nb, nc = G.ncart(cls[0]), G.ncart(cls[1]), G.ncart(cls[2])
        cover = np.zeros((na, nb, nc), dtype=int)
        for p in parts:
            (a0, a1), (b0, b1), (c0, c1) = p["a"], p["b"], p["c"]
            cover[a0:a1, b0:b1, c0:c1] += 1
            assert p["dinv"] and len(p["combos"]) == 9
        assert (cover == 1).all(), cls
        assert len(parts) <= 128


def test_sph_transform_orthonormal():
    for l in (2, 3):
        T = sph_transform(l)
        S = cart_overlap_shared_norm(l)
        assert np.abs(T @ S @ T.T - np.eye(2 * l + 1)).max() < 1e-14


def test_basis_sizes():
    m = workloads.cluster10()
    assert BasisSet.build(m, "even-tempered").nbf() == 300
    assert BasisSet.build(m, "even-tempered", pure=True).nbf() == 280
    assert BasisSet.build(workloads.cluster10_subset4(), "even-tempered", pure=True).nbf() == 112
    assert BasisSet.build(workloads.cluster24(), "even-tempered", pure=True).nbf() == 672
    w = workloads.water_sto3g()
    b = BasisSet.build(w, "sto-3g")
    assert b.nbf() == 7 and [s.l for s in b.shells] == [0, 0, 1, 0, 0]


def test_synthetic_geometries_match_fixture():
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "geometries.json")))
    for name, mol in [("cluster10", workloads.cluster10()), ("cluster24", workloads.cluster24())]:
        assert fx[name]["symbols"] == [mol.symbol(i) for i in range(mol.natom())]
        assert np.allclose(fx[name]["xyz_bohr"], mol.geometry(), atol=0, rtol=0)
        g = mol.geometry()
        d = np.linalg.norm(g[:, None] - g[None], axis=-1) + np.eye(len(g)) * 10
        assert d.min() >= 1.8 and g.min() >= 0 and g.max() < 8.0
    b = workloads.batch256()
    assert len(b) == 256 and len(fx["batch256"]) == 256
    for m, f in zip(b, fx["batch256"]):
        assert f["symbols"] == [m.symbol(i) for i in range(m.natom())]
        assert 3 <= m.natom() <= 12 and m.nelectron() % 2 == 0


def test_xyz_round_trip(tmp_path):
    m = workloads.cluster10()
    p = str(tmp_path / "g.xyz")
    m.save_xyz_file(p, True)
    syms, xyz = read_xyz(p)
    assert syms == [m.symbol(i) for i in range(m.natom())]
    assert np.abs(xyz - m.geometry()).max() < 1e-14


def test_fock_eigensolver_host_and_device_paths_agree():
    """The SCF's Fock diagonalisation (LAPACK on the host for small n) gives
    the eigenpairs torch.linalg.eigh gives."""
    import torch
    from quax_amd.methods.hartree_fock import _eigh
    rng = np.random.default_rng(3)
    X = rng.standard_normal((40, 40))
    F = torch.tensor(X + X.T, dtype=torch.float64)
    w1, v1 = _eigh(F, "host")
    w2, v2 = _eigh(F, "device")
    assert torch.allclose(w1, w2, atol=1e-12)
    # eigenvectors up to sign: compare projectors onto the lowest 10
    P1, P2 = v1[:, :10] @ v1[:, :10].T, v2[:, :10] @ v2[:, :10].T
    assert torch.allclose(P1, P2, atol=1e-10)


def test_fock_eigensolver_occupied_subset():
    """The occupied-only host solve (dsyevr) gives the lowest eigenvalues and
    the same occupied projector (the density) as the full solve."""
    import torch
    from quax_amd.methods.hartree_fock import _eigh
    rng = np.random.default_rng(5)
    X = rng.standard_normal((60, 60))
    F = torch.tensor(X + X.T, dtype=torch.float64)
    w, v = _eigh(F, "host")
    ws, vs = _eigh(F, "host", nocc=12)
    assert ws.shape == (12,) and vs.shape == (60, 12)
    assert torch.allclose(ws, w[:12], atol=1e-12)
    assert torch.allclose(vs @ vs.T, v[:, :12] @ v[:, :12].T, atol=1e-12)
    wf, vf = _eigh(F, "host", nocc=60)      # nocc = n falls back to the full solve
    assert wf.shape == (60,)


def test_tuned_tables_reach_the_generated_class_table(tmp_path):
    """The autotuner's JSON (quartet tiles, FP32 policy) becomes the default
    tables the runtime compiles in: qx_tile_default[cls] = {J/K, gradient}
    threads per block, qx_fp32_default[cls] = {J/K, gradient} flags."""
    import json
    import subprocess
    import sys
    tuning = {"tiles": {"jk": {"1110": 64}, "grad": {"1111": 256, "1000": 192}},
              "fp32": {"jk": ["1111"], "grad": ["1110", "1111"]}, "grad_wave_reduce": ["1100", "0000"]}
    f = tmp_path / "tuning.json"
    f.write_text(json.dumps(tuning))
    out = tmp_path / "gen"
    env = dict(os.environ, QX_TUNING_FILE=str(f))
    subprocess.run([sys.executable, os.path.join(ROOT, "quax_amd", "csrc", "gen_eri.py"), str(out), "1"],
                   check=True, capture_output=True, env=env)
    inc = (out / "qx_classes.inc").read_text()
    classes = re.findall(r"^  \{(\d), (\d), (\d), (\d), qx_launch_e", inc, re.M)
    names = ["".join(c) for c in classes]

    def table(name):
        body = inc.split(f"{name}[QX_NUM_CLASSES][2] = {{", 1)[1].split("};", 1)[0]
        return [tuple(int(x) for x in r) for r in re.findall(r"\{(\d+), (\d+)\}", body)]

    tiles, fp32 = table("qx_tile_default"), table("qx_fp32_default")
    assert len(tiles) == len(fp32) == len(names)
    gw = [int(x) for x in inc.split("qx_gwred_default[QX_NUM_CLASSES] = {", 1)[1].split("};", 1)[0].split(",")]
    assert gw == [int(nm in tuning["grad_wave_reduce"]) for nm in names]
    for nm, t, f32 in zip(names, tiles, fp32):
        assert t == (tuning["tiles"]["jk"].get(nm, 128), tuning["tiles"]["grad"].get(nm, 128)), nm
        assert f32 == (int(nm in tuning["fp32"]["jk"]), int(nm in tuning["fp32"]["grad"])), nm


def test_streaming_kernels_fit_side_stream_scratch():
    """Every streaming J/K and adjoint-gradient kernel of the last build needs
    at most 6 KB of scratch per lane (the side streams' budget; read from the
    compiled code objects), and the reader finds the kernels' sizes."""
    import glob
    from quax_amd.csrc import build
    objs = {}
    for f in glob.glob(os.path.join(build.OBJ, "cls_*_[jg]*.hip.*.o")):
        src = os.path.basename(f).split(".hip.")[0]
        if src not in objs or os.path.getmtime(f) > os.path.getmtime(objs[src]):
            objs[src] = f
    if not objs or not os.path.exists(os.path.join(build.LLVM_BIN, "clang-offload-bundler")):
        pytest.skip("no built objects or no offload bundler")
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        sizes = build.kernel_scratch(objs["cls_2221_j0"], tmp)
    assert sizes and all(name.startswith("_Z") and size >= 0 for name, size in sizes)
    build.check_streaming_scratch(list(objs.values()), verbose=False)
