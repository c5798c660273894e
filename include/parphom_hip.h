/*
 * parphom_hip.h - C ABI of the MI355X phonon library (libparphom_hip.so).
 *
 * Drop-in for the per-q hot path of parphom.x: dynamical-matrix assembly,
 * the pzheevx eigensolve and the analytic group velocity.  A host program
 * (the Fortran driver through ISO_C_BINDING, Python through ctypes) keeps
 * reading the input files and writing the HDF5 output; it hands the
 * geometry, the (N,N,3,3) force constants and the q-points to this library.
 *
 * Conventions
 *   - All integers are 64-bit; arrays are passed as plain pointers.
 *   - Two-dimensional arrays are column-major (Fortran order), so a Fortran
 *     caller passes its arrays unchanged.
 *   - Complex numbers are interleaved (re, im) doubles (Fortran
 *     double complex, C double _Complex).
 *   - Return codes: 0 success, < 0 error (text via pph_last_error).  Nothing
 *     in this library exits the process.
 *   - A handle is bound to one GPU and is not re-entrant; calls on one
 *     handle must not overlap.  pph_run* are synchronous at return.
 */
#ifndef PARPHOM_HIP_H
#define PARPHOM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PPH_OK 0
#define PPH_ERR_ARG (-1)
#define PPH_ERR_HIP (-2)
#define PPH_ERR_MEMORY (-3)
#define PPH_ERR_INTERNAL (-4)

/* per-q info bits, with pzheevx's meaning (diagonalize.F90:100-115) */
#define PPH_INFO_NOCONV 1     /* an eigenvalue failed to converge */
#define PPH_INFO_ORTH 2       /* eigenvectors of a cluster could not be orthogonalised */
#define PPH_INFO_MEMORY 4     /* not all eigenvectors fit the caller's evec_cols */
#define PPH_INFO_NOTCOMP 8    /* some eigenvalues were not computed */

typedef struct pph_handle pph_handle;

/* Geometry: the fields of parphom's `moire` state after read_lammps_data
 * (mod_variables.F90:83-89, lammps_read.F90:28-243). */
typedef struct pph_system {
  int64_t natom;          /* N */
  int64_t ntypes;
  const double* lat;      /* 3x3, lat(i,k) = component k of lattice vector i, at lat[i + 3k] */
  const double* pos;      /* N x 3 Cartesian positions (A), pos(i,k) at pos[i + N k] */
  const double* mass;     /* ntypes masses (amu) */
  const int64_t* type;    /* N atom types, 1-based */
} pph_system;

/* Options: the pzheevx_vars / velocity settings of read_input.F90:368-516. */
typedef struct pph_opts {
  char jobz;              /* 'N' eigenvalues only, 'V' with eigenvectors (forced by vel) */
  char range;             /* 'A' all, 'V' interval, 'I' index range */
  char vel_method;        /* 'A' analytic dD/dq, 'C' central difference */
  char precision;         /* 'D' FP64 throughout, 'S' FP32 matrix cores + FP64 refinement.
                             On 'S' the eigenvalues are FP64 Rayleigh quotients of the
                             FP32 vectors (refined); bisection of the FP32 tridiagonal,
                             which only supplies the inverse-iteration shifts, stops at
                             max(abstol, 1e-13 |T|); with range 'V' a mode whose refined
                             eigenvalue leaves (vl, vu] is dropped */
  int32_t vel;            /* 1: group velocities */
  int32_t device;         /* HIP device ordinal */
  double vl, vu;          /* range 'V' bounds in the input file's units (cm^-1); like
                             parphom.x they are divided by f32(15.633302*33.35641) and
                             compared with lambda = omega^2 (diagonalize.F90:73-74) */
  int64_t il, iu;         /* range 'I', 1-based inclusive */
  double abstol, orfac;   /* pzheevx ABSTOL / ORFAC */
  int64_t batch;          /* q-points per device batch, 0 = automatic */
  int32_t streams;        /* batches in flight in pph_run (one stream and host thread
                             each, so one batch's memory-bound reduction overlaps
                             another's matrix-core work); 0 = automatic (2 when the
                             q-points make at least two batches) */
} pph_opts;

/* Outputs of pph_run (host pointers) or pph_run_device (device pointers).
 * Per q-point iq (0-based):
 *   eval_cm[iq*ld_eval + k], k < m[iq]: signed frequencies in cm^-1, ascending
 *       (write_output.F90:135), the `eigenvalues` dataset;
 *   evec (optional): complex n x nz[iq] column-major at evec + iq*n*evec_cols,
 *       the bytes of the HDF5 `evec` dataset (h5py shape (nz, n));
 *   vel (optional): vel[(iq*evec_cols + k)*3 + a] in m/s, the `vel` dataset
 *       (h5py shape (nz, 3), write_output.F90:397-446);
 *   info[iq]: PPH_INFO_* bits. */
typedef struct pph_out {
  double* eval_cm;
  int64_t ld_eval;
  int64_t* m;
  int64_t* nz;
  void* evec;
  int64_t evec_cols;
  double* vel;
  int32_t* info;
} pph_out;

/* Build the device state: symmetrised force constants (force_read.F90:187-189,
 * replaces read_force_constants + pdgeadd), crystal coordinates and reciprocal
 * lattice (lammps_read.F90:198-228), minimum-image tables.
 * fc: C-order (N,N,3,3) float64, exactly the HDF5 `force_constants` dataset
 * (a Fortran caller's fc(3,3,N,N) array has the same bytes). */
int pph_create(const pph_system* sys, const double* fc, const pph_opts* opts, pph_handle** h);

/* The q loop of calculations.F90:49-53 (create_dynamical_matrix,
 * diagonalize_dynamical_matrix, the velocity block of write_output.F90:410-466)
 * for nq q-points in reduced coordinates, q(iq,k) at q[iq + ldq k]
 * (q_file%points). Outputs are host pointers. */
int pph_run(pph_handle* h, int64_t nq, const double* q, int64_t ldq, pph_out* out);

/* As pph_run, with every output pointer in `out` a device pointer on the
 * handle's GPU (q stays a host pointer). */
int pph_run_device(pph_handle* h, int64_t nq, const double* q, int64_t ldq, pph_out* out);

/* Dynamical matrices alone (create_dynamical_matrix, dynamical_matrix.F90:49-191).
 * what: 0 = D(q); 1, 2, 3 = dD/dq_x,y,z with the handle's vel_method.
 * out: complex128 n x n column-major per q at out + iq*n*n (host or device
 * pointer per `on_device`). */
int pph_assemble(pph_handle* h, int64_t nq, const double* q, int64_t ldq, int32_t what, void* out,
                 int32_t on_device);

/* Expectation values diag(Z^H X Z) for nq q-points, with X = D(q) (what 0)
 * or dD/dq_x,y,z (what 1..3) assembled from this handle's force constants:
 * out[iq*nz + k] = Re(z_k^H X(q) z_k), Z complex128 n x nz per q at
 * Z + iq*n*zcols (column-major).  With a handle built on the derivative of
 * the force constants with respect to a model parameter this is the
 * Hellmann-Feynman derivative d lambda_k / d parameter; with what = 1..3 it
 * is <e|dD/dq|e> of write_output.F90:415-421.  Pointers are host or device
 * per `on_device`. */
int pph_expect(pph_handle* h, int64_t nq, const double* q, int64_t ldq, int32_t what,
               const void* Z, int64_t nz, int64_t zcols, double* out, int32_t on_device);

/* The eigensolver alone (pzheevx semantics, diagonalize.F90:73-96) on nmat
 * Hermitian n x n complex128 matrices (column-major, leading dimension n,
 * consecutive in memory; destroyed).  Precision, jobz, range, vl, vu
 * (already in lambda units here), il, iu, abstol and orfac come from opts.
 * w[imat*n + k] eigenvalues (lambda, ascending), Z optional n x n per matrix,
 * m[imat] count found, info[imat] bits.  All pointers are host or device
 * per `on_device`. */
int pph_heevx(int64_t n, int64_t nmat, void* A, const pph_opts* opts, double* w, void* Z,
              int64_t* m, int32_t* info, int32_t on_device);

/* Density of states by linear triangulation: the loops of
 * bz_integration.f90:12-47 (linear_triangulation) with sort_freq (:49-64) and
 * integrate_triangle (:67-134), the hot part of
 * moire_phonon_utils.density_of_states (pyphutil.py:227-306).
 * freq, func: nk x nbands row-major (freq[k*nbands + b], the `eigenvalues`
 * of each q group, cm^-1); func may be NULL for 1 everywhere (the
 * reference's default, pyphutil.py:235-236); tri: ntri x 3 row-major,
 * 0-based indices into the nk rows; E: nE energies, ascending.
 * dos[i], nd[i]: the reference's dos_loc and ndos_loc summed over all
 * triangles (before the 1/(3 ntri)/area factors of pyphutil.py:301-302).
 * Host or device pointers per `on_device`; the call is synchronous. */
int pph_dos_triangulation(int64_t nk, int64_t nbands, const double* freq, const double* func,
                          int64_t ntri, const int64_t* tri, int64_t nE, const double* E,
                          double* dos, double* nd, int32_t device, int32_t on_device);

/* pph_dos_triangulation without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same sums as pph_dos_triangulation; every pointer is a device
 * pointer on `device` (no host copies, no argument checks that would read the
 * arrays: vertex indices must lie in [0, nk) and E must be ascending), func
 * may be NULL (1 everywhere).  The kernels run asynchronously on `stream`
 * (a hipStream_t; NULL = the null stream) and the call returns after
 * enqueuing them.  The difference array of the triangulation is a static
 * device array of 2^20 doubles per GPU (8 MiB, held until the process ends):
 * nE above 2^20 is rejected with PPH_ERR_ARG.  Every call on one GPU shares
 * that array (the call clears it first), so calls whose execution overlaps on
 * one GPU (different streams, concurrent graph replays) race. */
int pph_dos_triangulation_noalloc(int64_t nk, int64_t nbands, const double* freq,
                                  const double* func, int64_t ntri, const int64_t* tri, int64_t nE,
                                  const double* E, double* dos, double* nd, int32_t device,
                                  void* stream);

/* Density of states by smearing, bz_integration.f90:140-199
 * (standard_integration with gaussian / lorentzian) summed over q-points as
 * pyphutil.py:311-325 does: dos[i] = sum_q weight[q]/natom sum_b
 * g(E[i] - freq[q*nbands + b]), method 1 Gaussian, 2 Lorentzian of `width`.
 * Each q-point's frequencies must be ascending (as parphom.x writes them).
 * Note: pyphutil passes its array of local weights where the Fortran expects
 * one number, so each MPI rank of the reference applies the weight of its
 * first q-point to all of its q-points; this entry applies weight[q] to
 * q-point q (the same result for uniform weights). */
int pph_dos_smearing(int32_t method, int64_t nq, int64_t nbands, const double* freq,
                     const double* weight, double width, int64_t natom, int64_t nE,
                     const double* E, double* dos, int32_t device, int32_t on_device);

/* Force-constant symmetrisation, the iterations of symmetrize_fc
 * (Force_Constant_Generation/get_force_constant.py:298-348): `iterations`
 * times, the acoustic sum rule (:320-331, fc[i, :, a, b] -= its mean over
 * the second atom) and then inversion symmetry (:334-348, fc[i, j, a, b] and
 * fc[j, i, b, a] both set to their mean), in place on the C-order
 * (natom, natom, 3, 3) float64 array (the HDF5 `force_constants` dataset).
 * Host or device pointer per `on_device`; synchronous. */
int pph_fc_symmetrize(int64_t natom, double* fc, int64_t iterations, int32_t device,
                      int32_t on_device);

/* Spring-model force constants on the GPU, the synthetic producer that
 * stands in for the finite-displacement LAMMPS producer of
 * Force_Constant_Generation/get_force_constant.py:227-295 (which writes the
 * dataset force_read.F90:93-158 reads).  For each listed pair p (atoms
 * pi[p] != pj[p], minimum-image separation rvec[3p .. 3p+2] in A) of kind
 * 0..2 (intralayer neighbour shell s: radial spring params[s], tangential
 * params[3 + s]) or 3 (interlayer: params[6] exp(-(|r| - inter_r0) /
 * inter_decay), radial):
 *   fc[i, j] = -(spring block) .* (1 + noise_amp u_ij),
 * u_ij symmetric in [-1, 1), drawn from (seed, pair) by a counter-based hash;
 * then fc[i, i] = -sum_{j != i} fc[i, j] (acoustic sum rule).  fc: dense
 * (natom, natom, 3, 3) C-order float64, eV/A^2.  The model is linear in
 * params, so params = e_k gives d fc / d params[k].  Host or device pointers
 * per `on_device` (with device pointers the pairs are not range-checked);
 * synchronous.  parphom_amd/synth.py builds the pair list and states the
 * same model on the CPU. */
int pph_spring_fc(int64_t natom, int64_t npairs, const int64_t* pi, const int64_t* pj,
                  const double* rvec, const int32_t* kind, const double* params, uint64_t seed,
                  double noise_amp, double inter_r0, double inter_decay, double* fc,
                  int32_t device, int32_t on_device);

int pph_destroy(pph_handle* h);

/* Last error text for a handle (or the thread's last error when h is NULL). */
const char* pph_last_error(const pph_handle* h);

/* Library version string. */
const char* pph_version(void);

#ifdef __cplusplus
}
#endif
#endif /* PARPHOM_HIP_H */
