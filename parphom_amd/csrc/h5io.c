/*
 * h5io.c - HDF5 input/output in parphom.x's layout (libparphom_h5.so).
 *
 * Reading: the (N, N, 3, 3) float64 force-constant dataset that
 * force_read.F90:93-158 reads (any file name, dataset name from the input).
 * Writing: one group per q-point named "%012d" (the 1-based q index,
 * write_output.F90:98) holding
 *   q_vec        float64 (3)                       write_output.F90:113-118
 *   eigenvalues  float64 (M), cm^-1                write_output.F90:134-140
 *   evec         compound {r: f64, i: f64} (NZ, n) write_output.F90:227-235
 *   vel          float64 (NZ, 3), m/s              write_output.F90:397-401
 * (shapes as h5py reports them; the Fortran writer declares the reversed
 * dimensions).  Serial HDF5; errors are returned as negative codes.
 */
#include <hdf5.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static char g_err[512];

static int fail(const char* msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return -1;
}

const char* pph_h5_last_error(void) { return g_err; }

static void quiet(void) { H5Eset_auto2(H5E_DEFAULT, NULL, NULL); }

int pph_h5_read_fc(const char* path, const char* dset, int64_t natom, double* out) {
  quiet();
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return fail("cannot open force-constant file");
  hid_t d = H5Dopen2(f, dset, H5P_DEFAULT);
  if (d < 0) { H5Fclose(f); return fail("force-constant dataset not found"); }
  hid_t s = H5Dget_space(d);
  hsize_t dims[8];
  int nd = H5Sget_simple_extent_ndims(s);
  int rc = 0;
  if (nd != 4) rc = fail("force-constant dataset must have rank 4");
  else {
    H5Sget_simple_extent_dims(s, dims, NULL);
    if ((int64_t)dims[0] != natom || (int64_t)dims[1] != natom || dims[2] != 3 || dims[3] != 3)
      rc = fail("force-constant dataset shape is not (natom, natom, 3, 3)");
    else if (H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
      rc = fail("reading the force-constant dataset failed");
  }
  H5Sclose(s);
  H5Dclose(d);
  H5Fclose(f);
  return rc;
}

int pph_h5_write_fc(const char* path, const char* dset, int64_t natom, const double* fc) {
  quiet();
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) return fail("cannot create force-constant file");
  hsize_t dims[4] = {(hsize_t)natom, (hsize_t)natom, 3, 3};
  hid_t s = H5Screate_simple(4, dims, NULL);
  hid_t d = H5Dcreate2(f, dset, H5T_IEEE_F64LE, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  int rc = 0;
  if (d < 0 || H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, fc) < 0)
    rc = fail("writing the force-constant dataset failed");
  if (d >= 0) H5Dclose(d);
  H5Sclose(s);
  H5Fclose(f);
  return rc;
}

/* overwrite the (natom, natom, 3, 3) dataset of an existing file in place */
int pph_h5_update_fc(const char* path, const char* dset, int64_t natom, const double* fc) {
  quiet();
  hid_t f = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
  if (f < 0) return fail("cannot open force-constant file for writing");
  hid_t d = H5Dopen2(f, dset, H5P_DEFAULT);
  if (d < 0) { H5Fclose(f); return fail("force-constant dataset not found"); }
  hid_t s = H5Dget_space(d);
  hsize_t dims[8];
  int rc = 0;
  if (H5Sget_simple_extent_ndims(s) != 4) rc = fail("force-constant dataset must have rank 4");
  else {
    H5Sget_simple_extent_dims(s, dims, NULL);
    if ((int64_t)dims[0] != natom || (int64_t)dims[1] != natom || dims[2] != 3 || dims[3] != 3)
      rc = fail("force-constant dataset shape is not (natom, natom, 3, 3)");
    else if (H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, fc) < 0)
      rc = fail("writing the force-constant dataset failed");
  }
  H5Sclose(s);
  H5Dclose(d);
  H5Fclose(f);
  return rc;
}

/* create (truncate) or open an output file; returns the file id in *fid */
int pph_h5_open_output(const char* path, int32_t create, int64_t* fid) {
  quiet();
  hid_t f = create ? H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
                   : H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
  if (f < 0) return fail("cannot open output file");
  *fid = (int64_t)f;
  return 0;
}

int pph_h5_open_readonly(const char* path, int64_t* fid) {
  quiet();
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return fail("cannot open file");
  *fid = (int64_t)f;
  return 0;
}

int pph_h5_close(int64_t fid) { return H5Fclose((hid_t)fid) < 0 ? fail("close failed") : 0; }

static int write_f64(hid_t g, const char* name, int rank, const hsize_t* dims, const double* data) {
  hid_t s = H5Screate_simple(rank, dims, NULL);
  hid_t d = H5Dcreate2(g, name, H5T_IEEE_F64LE, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  int rc = 0;
  if (d < 0 || H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) rc = -1;
  if (d >= 0) H5Dclose(d);
  H5Sclose(s);
  return rc;
}

int pph_h5_write_group(int64_t fid, int64_t qindex, const double* qvec, const double* eval_cm,
                       int64_t m, const double* evec, int64_t n, int64_t nz, const double* vel) {
  char name[32];
  snprintf(name, sizeof name, "%012lld", (long long)qindex);
  hid_t g = H5Gcreate2((hid_t)fid, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (g < 0) return fail("cannot create q group (exists already?)");
  int rc = 0;
  hsize_t d3[1] = {3};
  if (write_f64(g, "q_vec", 1, d3, qvec) < 0) rc = fail("writing q_vec failed");
  hsize_t dm[1] = {(hsize_t)m};
  if (!rc && write_f64(g, "eigenvalues", 1, dm, eval_cm) < 0) rc = fail("writing eigenvalues failed");
  if (!rc && evec && nz > 0) {
    hid_t ft = H5Tcreate(H5T_COMPOUND, 16);
    H5Tinsert(ft, "r", 0, H5T_IEEE_F64LE);
    H5Tinsert(ft, "i", 8, H5T_IEEE_F64LE);
    hid_t mt = H5Tcreate(H5T_COMPOUND, 16);
    H5Tinsert(mt, "r", 0, H5T_NATIVE_DOUBLE);
    H5Tinsert(mt, "i", 8, H5T_NATIVE_DOUBLE);
    hsize_t de[2] = {(hsize_t)nz, (hsize_t)n};
    hid_t s = H5Screate_simple(2, de, NULL);
    hid_t d = H5Dcreate2(g, "evec", ft, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (d < 0 || H5Dwrite(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, evec) < 0) rc = fail("writing evec failed");
    if (d >= 0) H5Dclose(d);
    H5Sclose(s);
    H5Tclose(mt);
    H5Tclose(ft);
  }
  if (!rc && vel && nz > 0) {
    hsize_t dv[2] = {(hsize_t)nz, 3};
    if (write_f64(g, "vel", 2, dv, vel) < 0) rc = fail("writing vel failed");
  }
  H5Gclose(g);
  return rc;
}

/* Read a float64 (or {r,i} compound, as interleaved doubles) dataset at
 * `path` inside an open file.  dims (up to 4) and rank are returned; the data
 * is copied when out != NULL and cap (doubles) is large enough. */
int pph_h5_read_dataset(int64_t fid, const char* path, double* out, int64_t cap, int64_t* dims,
                        int32_t* rank, int32_t* is_complex) {
  hid_t d = H5Dopen2((hid_t)fid, path, H5P_DEFAULT);
  if (d < 0) return fail("dataset not found");
  hid_t s = H5Dget_space(d);
  hid_t t = H5Dget_type(d);
  hsize_t hd[8] = {0};
  int nd = H5Sget_simple_extent_ndims(s);
  int rc = 0;
  if (nd > 4) rc = fail("rank > 4");
  else {
    H5Sget_simple_extent_dims(s, hd, NULL);
    int64_t count = 1;
    for (int i = 0; i < nd; ++i) { dims[i] = (int64_t)hd[i]; count *= (int64_t)hd[i]; }
    *rank = nd;
    const int cpx = (H5Tget_class(t) == H5T_COMPOUND);
    *is_complex = cpx;
    if (cpx) count *= 2;
    if (out) {
      if (count > cap) rc = fail("buffer too small");
      else if (cpx) {
        hid_t mt = H5Tcreate(H5T_COMPOUND, 16);
        H5Tinsert(mt, "r", 0, H5T_NATIVE_DOUBLE);
        H5Tinsert(mt, "i", 8, H5T_NATIVE_DOUBLE);
        if (H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0) rc = fail("read failed");
        H5Tclose(mt);
      } else if (H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0) {
        rc = fail("read failed");
      }
    }
  }
  H5Tclose(t);
  H5Sclose(s);
  H5Dclose(d);
  return rc;
}

/* number of top-level groups, and the name of group `idx` */
int pph_h5_group_count(int64_t fid, int64_t* count) {
  H5G_info_t info;
  if (H5Gget_info((hid_t)fid, &info) < 0) return fail("group info failed");
  *count = (int64_t)info.nlinks;
  return 0;
}

int pph_h5_group_name(int64_t fid, int64_t idx, char* name, int64_t cap) {
  ssize_t len = H5Lget_name_by_idx((hid_t)fid, ".", H5_INDEX_NAME, H5_ITER_INC, (hsize_t)idx, name,
                                   (size_t)cap, H5P_DEFAULT);
  return len < 0 ? fail("group name failed") : 0;
}
