/*
 * parphom_h5.h - HDF5 input/output in parphom.x's layout (libparphom_h5.so,
 * serial libhdf5).  Replaces, for a host without parallel HDF5:
 *   pph_h5_read_fc      read_force_constants' dataset read (force_read.F90:93-158)
 *   pph_h5_write_group  the per-q group of write_output (write_output.F90:95-464):
 *                       "%012d" / q_vec (3), eigenvalues (M), evec {r,i} (NZ, n),
 *                       vel (NZ, 3)   [h5py shapes]
 * The remaining entries open/close files and read datasets back (tests,
 * post-processing).  Return 0 on success, -1 on error (pph_h5_last_error).
 */
#ifndef PARPHOM_H5_H
#define PARPHOM_H5_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
int pph_h5_read_fc(const char* path, const char* dset, int64_t natom, double* out);
int pph_h5_write_fc(const char* path, const char* dset, int64_t natom, const double* fc);
/* overwrite the existing (natom, natom, 3, 3) dataset in place (the in-place
 * update symmetrize_fc makes, get_force_constant.py:307-317) */
int pph_h5_update_fc(const char* path, const char* dset, int64_t natom, const double* fc);
int pph_h5_open_output(const char* path, int32_t create, int64_t* fid);
int pph_h5_open_readonly(const char* path, int64_t* fid);
int pph_h5_write_group(int64_t fid, int64_t qindex, const double* qvec, const double* eval_cm,
                       int64_t m, const double* evec, int64_t n, int64_t nz, const double* vel);
int pph_h5_read_dataset(int64_t fid, const char* path, double* out, int64_t cap, int64_t* dims,
                        int32_t* rank, int32_t* is_complex);
int pph_h5_group_count(int64_t fid, int64_t* count);
int pph_h5_group_name(int64_t fid, int64_t idx, char* name, int64_t cap);
int pph_h5_close(int64_t fid);
const char* pph_h5_last_error(void);
#ifdef __cplusplus
}
#endif
#endif
