/* quax_amd_h5.h -- HDF5 files of ERI derivative tensors for Quax's disk mode
 * (libquax_h5.so, host code on the HDF5 C library).
 *
 * Replaces the HDF5 half of compute_2e_deriv_disk (quax/integrals/
 * libint_interface.cc:1811-2115: H5File, one DataSet per derivative order,
 * hyperslab writes) and the h5py reads of TEI's disk mode (quax/integrals/
 * tei.py:156-181) and check_tei_disk (quax/methods/ints.py:331-360).  The
 * tensors themselves are computed on the GPU (qx_eri_deriv_core) by the
 * caller, which hands host blocks to qx_h5_write_block.
 *
 * Every call returns 0 or a nonzero code (1 bad argument, 2 HDF5 I/O error,
 * 3 shape mismatch, 4 no such dataset); qx_h5_last_error() explains. */
#ifndef QUAX_AMD_H5_H
#define QUAX_AMD_H5_H

#ifdef __cplusplus
extern "C" {
#endif

#define QX_H5_MAX_RANK 8
enum { QX_H5_READ = 0, QX_H5_TRUNCATE = 1, QX_H5_APPEND = 2 };

/* Writes rows [first, first + count) along dimension 0 of the FP64 dataset
 * dset (rank <= 8, shape dims; created zero-filled if absent, else its shape
 * must equal dims) from the row-major host block `data`.  mode: TRUNCATE
 * starts a new file, APPEND opens (or creates) one. */
int qx_h5_write_block(const char* path, int mode, const char* dset, int rank, const long long* dims,
                      long long first, long long count, const double* data);

/* Rank and dimensions of a dataset (dims holds QX_H5_MAX_RANK entries). */
int qx_h5_dataset_dims(const char* path, const char* dset, int* rank, long long* dims);

/* Rank-5 dataset: the slice [:, :, :, :, index] (the layout of
 * <type>_derivs.h5); any other rank: the whole dataset (the 4-D datasets of
 * <type>_partials.h5).  out: host buffer of `capacity` doubles. */
int qx_h5_read_slice(const char* path, const char* dset, long long index, double* out, long long capacity);

const char* qx_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
