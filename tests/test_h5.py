"""HDF5 files of the disk mode (libquax_h5.so, include/quax_amd_h5.h) on the
CPU: blocks written along the first dimension read back as the reference's
disk mode reads them (tei.py:173-180: [:, :, :, :, idx] of a rank-5 dataset,
the whole of a rank-4 one), the header checked with HDF5's own h5dump, and
the error paths."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from quax_amd import _h5

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = _h5.load()
    header = open(os.path.join(ROOT, "include", "quax_amd_h5.h")).read()
    names = sorted(set(re.findall(r"\b(qx_h5_[a-z0-9_]+)\s*\(", header)))
    assert len(names) == 4
    for nm in names:
        assert hasattr(L, nm), nm


def test_blocks_round_trip_and_slices(tmp_path):
    rng = np.random.default_rng(1)
    n, nd = 5, 7
    T = rng.standard_normal((n, n, n, n, nd))
    path = str(tmp_path / "eri_derivs.h5")
    name = f"eri_{n}_{n}_{n}_{n}_deriv1"
    # rows written out of order, in blocks of one and two
    _h5.write_block(path, name, T.shape, 3, T[3:5], mode=_h5.TRUNCATE)
    _h5.write_block(path, name, T.shape, 0, T[0:1])
    _h5.write_block(path, name, T.shape, 1, T[1:3])
    assert _h5.dataset_dims(path, name) == T.shape
    for idx in (0, 3, nd - 1):
        assert np.array_equal(_h5.read_slice(path, name, idx), T[..., idx])
    # a second dataset in the same file; a 4-D one reads whole
    U = rng.standard_normal((n, n, n, n))
    _h5.write_block(path, "eri_5_5_5_5_deriv2_4", U.shape, 0, U)
    assert np.array_equal(_h5.read_slice(path, "eri_5_5_5_5_deriv2_4"), U)
    assert np.array_equal(_h5.read_slice(path, name, 2), T[..., 2])
    # rows never written read as the fill value 0
    _h5.write_block(path, "eri_5_5_5_5_deriv2", (n, n, n, n, 3), 0, np.ones((1, n, n, n, 3)))
    part = _h5.read_slice(path, "eri_5_5_5_5_deriv2", 1)
    assert (part[0] == 1.0).all() and (part[1:] == 0.0).all()


@pytest.mark.skipif(not os.path.exists("/opt/conda/bin/h5dump"), reason="h5dump not installed")
def test_header_as_hdf5_reads_it(tmp_path):
    path = str(tmp_path / "eri_derivs.h5")
    T = np.arange(2 * 2 * 2 * 2 * 3, dtype=np.float64).reshape(2, 2, 2, 2, 3)
    _h5.write_block(path, "eri_2_2_2_2_deriv1", T.shape, 0, T, mode=_h5.TRUNCATE)
    out = subprocess.run(["/opt/conda/bin/h5dump", "-H", path], capture_output=True, text=True, check=True).stdout
    assert 'DATASET "eri_2_2_2_2_deriv1"' in out
    assert "H5T_IEEE_F64LE" in out
    assert "( 2, 2, 2, 2, 3 )" in out
    vals = subprocess.run(["/opt/conda/bin/h5dump", "-d", "eri_2_2_2_2_deriv1", "-w", "0", path],
                          capture_output=True, text=True, check=True).stdout
    nums = [float(x) for x in re.findall(r"-?\d+(?:\.\d+)?", vals.split("DATA {", 1)[1].split("}", 1)[0])
            if "(" not in x]
    # h5dump prints "(i,j,k,l,m): v, ..." rows; every stored value appears in row-major order
    assert sorted(set(nums) & set(T.reshape(-1).tolist())) == sorted(set(T.reshape(-1).tolist()))


def test_errors(tmp_path):
    path = str(tmp_path / "x.h5")
    assert _h5.dataset_dims(path, "nothing") is None
    _h5.write_block(path, "a", (2, 3), 0, np.zeros((2, 3)), mode=_h5.TRUNCATE)
    assert _h5.dataset_dims(path, "b") is None
    with pytest.raises(_h5.H5Error):
        _h5.write_block(path, "a", (2, 4), 0, np.zeros((2, 4)))    # shape differs from the existing dataset
    with pytest.raises(_h5.H5Error):
        _h5.write_block(path, "a", (2, 3), 1, np.zeros((2, 3)))    # rows beyond the first dimension
    with pytest.raises(_h5.H5Error):
        _h5.read_slice(path, "missing")
    with pytest.raises(ValueError):
        _h5.write_block(path, "a", (2, 3), 0, np.zeros((1, 4)))
