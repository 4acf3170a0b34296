/* HDF5 files of integral derivative tensors in the layout Quax's disk mode
 * uses (compute_2e_deriv_disk writes them, libint_interface.cc:1811-2115;
 * TEI's disk mode reads them, tei.py:156-181; check_tei_disk inspects them,
 * ints.py:331-360): one FP64 dataset per derivative order named
 * "<type>_<n1>_<n2>_<n3>_<n4>_deriv<order>" of shape (n1, n2, n3, n4,
 * nderivs), zero fill, or 4-D "..._deriv<order>_<index>" datasets in the
 * partials file.  Host code only (HDF5 C API); the tensors come from the GPU
 * through the caller. */
#include <hdf5.h>
#include <stdio.h>
#include <string.h>

#include "quax_amd_h5.h"

static char g_err[512];

static int fail(int code, const char* msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return code;
}

const char* qx_h5_last_error(void) { return g_err; }

static void quiet(void) { H5Eset_auto2(H5E_DEFAULT, NULL, NULL); }

static hid_t open_file(const char* path, int mode) {
  if (mode == QX_H5_TRUNCATE) return H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  hid_t f = H5Fopen(path, mode == QX_H5_READ ? H5F_ACC_RDONLY : H5F_ACC_RDWR, H5P_DEFAULT);
  if (f < 0 && mode == QX_H5_APPEND) f = H5Fcreate(path, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
  return f;
}

int qx_h5_write_block(const char* path, int mode, const char* dset, int rank, const long long* dims,
                      long long first, long long count, const double* data) {
  if (!path || !dset || !dims || rank < 1 || rank > QX_H5_MAX_RANK || (count > 0 && !data) || first < 0 ||
      count < 0 || first + count > dims[0] || (mode != QX_H5_TRUNCATE && mode != QX_H5_APPEND))
    return fail(1, "qx_h5_write_block: bad argument");
  quiet();
  hsize_t d[QX_H5_MAX_RANK], start[QX_H5_MAX_RANK], cnt[QX_H5_MAX_RANK];
  for (int k = 0; k < rank; ++k) {
    if (dims[k] < 0) return fail(1, "qx_h5_write_block: negative dimension");
    d[k] = (hsize_t)dims[k];
    start[k] = 0;
    cnt[k] = d[k];
  }
  start[0] = (hsize_t)first;
  cnt[0] = (hsize_t)count;
  int rc = 0;
  hid_t f = open_file(path, mode);
  if (f < 0) return fail(2, "qx_h5_write_block: cannot open or create the file");
  hid_t ds = -1, fs = -1, ms = -1, pl = -1;
  if (H5Lexists(f, dset, H5P_DEFAULT) > 0) {
    ds = H5Dopen2(f, dset, H5P_DEFAULT);
    fs = ds >= 0 ? H5Dget_space(ds) : -1;
    hsize_t have[QX_H5_MAX_RANK];
    if (fs < 0 || H5Sget_simple_extent_ndims(fs) != rank || H5Sget_simple_extent_dims(fs, have, NULL) != rank) {
      rc = fail(3, "qx_h5_write_block: existing dataset has another rank");
      goto done;
    }
    for (int k = 0; k < rank; ++k)
      if (have[k] != d[k]) {
        rc = fail(3, "qx_h5_write_block: existing dataset has another shape");
        goto done;
      }
  } else {
    fs = H5Screate_simple(rank, d, NULL);
    pl = H5Pcreate(H5P_DATASET_CREATE);
    const double zero = 0.0;
    H5Pset_fill_value(pl, H5T_NATIVE_DOUBLE, &zero);
    ds = H5Dcreate2(f, dset, H5T_NATIVE_DOUBLE, fs, H5P_DEFAULT, pl, H5P_DEFAULT);
    if (ds < 0) {
      rc = fail(2, "qx_h5_write_block: cannot create the dataset");
      goto done;
    }
  }
  if (count > 0) {
    ms = H5Screate_simple(rank, cnt, NULL);
    if (H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, NULL, cnt, NULL) < 0 ||
        H5Dwrite(ds, H5T_NATIVE_DOUBLE, ms, fs, H5P_DEFAULT, data) < 0)
      rc = fail(2, "qx_h5_write_block: write failed");
  }
done:
  if (ms >= 0) H5Sclose(ms);
  if (pl >= 0) H5Pclose(pl);
  if (fs >= 0) H5Sclose(fs);
  if (ds >= 0) H5Dclose(ds);
  H5Fclose(f);
  return rc;
}

int qx_h5_dataset_dims(const char* path, const char* dset, int* rank, long long* dims) {
  if (!path || !dset || !rank || !dims) return fail(1, "qx_h5_dataset_dims: bad argument");
  quiet();
  hid_t f = open_file(path, QX_H5_READ);
  if (f < 0) return fail(2, "qx_h5_dataset_dims: cannot open the file");
  int rc = 0;
  if (H5Lexists(f, dset, H5P_DEFAULT) <= 0) {
    H5Fclose(f);
    return fail(4, "qx_h5_dataset_dims: no such dataset");
  }
  hid_t ds = H5Dopen2(f, dset, H5P_DEFAULT);
  hid_t fs = ds >= 0 ? H5Dget_space(ds) : -1;
  const int r = fs >= 0 ? H5Sget_simple_extent_ndims(fs) : -1;
  if (r < 1 || r > QX_H5_MAX_RANK) {
    rc = fail(3, "qx_h5_dataset_dims: unsupported rank");
  } else {
    hsize_t d[QX_H5_MAX_RANK];
    H5Sget_simple_extent_dims(fs, d, NULL);
    *rank = r;
    for (int k = 0; k < r; ++k) dims[k] = (long long)d[k];
  }
  if (fs >= 0) H5Sclose(fs);
  if (ds >= 0) H5Dclose(ds);
  H5Fclose(f);
  return rc;
}

int qx_h5_read_slice(const char* path, const char* dset, long long index, double* out, long long capacity) {
  if (!path || !dset || !out) return fail(1, "qx_h5_read_slice: bad argument");
  int rank = 0;
  long long dims[QX_H5_MAX_RANK];
  int rc = qx_h5_dataset_dims(path, dset, &rank, dims);
  if (rc) return rc;
  long long n = 1;
  const int keep = rank == 5 ? 4 : rank;  /* rank 5: [..., index]; otherwise the whole dataset */
  for (int k = 0; k < keep; ++k) n *= dims[k];
  if (rank == 5 && (index < 0 || index >= dims[4])) return fail(1, "qx_h5_read_slice: index out of range");
  if (n > capacity) return fail(1, "qx_h5_read_slice: output buffer too small");
  quiet();
  hid_t f = open_file(path, QX_H5_READ);
  if (f < 0) return fail(2, "qx_h5_read_slice: cannot open the file");
  hid_t ds = H5Dopen2(f, dset, H5P_DEFAULT);
  hid_t fs = ds >= 0 ? H5Dget_space(ds) : -1;
  hsize_t start[QX_H5_MAX_RANK], cnt[QX_H5_MAX_RANK], mdim[1] = {(hsize_t)n};
  for (int k = 0; k < rank; ++k) {
    start[k] = 0;
    cnt[k] = (hsize_t)dims[k];
  }
  if (rank == 5) {
    start[4] = (hsize_t)index;
    cnt[4] = 1;
  }
  hid_t ms = H5Screate_simple(1, mdim, NULL);
  if (fs < 0 || H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, NULL, cnt, NULL) < 0 ||
      H5Dread(ds, H5T_NATIVE_DOUBLE, ms, fs, H5P_DEFAULT, out) < 0)
    rc = fail(2, "qx_h5_read_slice: read failed");
  H5Sclose(ms);
  if (fs >= 0) H5Sclose(fs);
  if (ds >= 0) H5Dclose(ds);
  H5Fclose(f);
  return rc;
}
