/*
 * parphom_hip_debug.h - test hook of libparphom_hip.so (not part of the
 * parphom.x drop-in surface; no counterpart in the reference).
 *
 * pph_debug_zgemm runs the library's matrix-core complex GEMM once on host
 * data so the GEMM can be checked in isolation:
 *   mode 0: C = op(A) op(B)
 *   mode 1: dot[j] = sum_i Re(conj(C_in[i,j]) (op(A) op(B))[i,j]), C unchanged
 *   mode 2: C = op(A) op(B) through the split-K path (4 pieces)
 *   mode 3: C = op(A) op(B) through the two-product path (opA = opB = 'N')
 *   mode 4: as mode 1 for Hermitian A (m = k, opA = 'N'), reading only the
 *           lower triangle of A (the half-work path of the velocity and
 *           Rayleigh-quotient products)
 *   mode 5: C = A B for Hermitian A (m = k, opA = 'N') from the lower
 *           triangle of A alone (the two triangular products of the band
 *           reduction's panel update, with four real products per complex one)
 *   mode 6: C = op(A) op(B) on the BF16 matrix cores (prec 'S' only: the
 *           complex64 inputs rounded to BF16, FP32 accumulation; the first
 *           step of the vector correction of the FP32 path)
 * prec 'D' complex128, 'S' complex64; column-major; returns 0 or < 0.
 */
#ifndef PARPHOM_HIP_DEBUG_H
#define PARPHOM_HIP_DEBUG_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
int pph_debug_zgemm(char prec, char opA, char opB, int64_t m, int64_t n, int64_t k,
                    const void* A, int64_t lda, const void* B, int64_t ldb, void* C,
                    int64_t ldc, int32_t mode, double* dot);
/* Phase timers measured with HIP events on the handle's stream around each
 * phase of pph_run / pph_run_device.  Slots: 0 assembly, 1 tridiagonal
 * reduction, 2 tridiagonal eigensolver, 3 back-transformation, 4 FP64
 * refinement, 5 velocity products, 6 whole batch, 7 the k_hemv_sym launches
 * of the one-stage reduction alone, sampled at columns j = 0, 8, 16, ... (an
 * event pair around each sampled launch; the count is the number of launches
 * timed), and for the two-stage reduction 8 dense -> band, 9 bulge chasing
 * (both inside slot 1) and 10 the k_q2_apply launch that applies the
 * bulge-chasing reflectors (inside slot 3).  ms/counts (11 entries each, may
 * be NULL) receive the accumulated milliseconds and counts; reset != 0 clears
 * them; enable turns recording on or off. */
int pph_profile(void* h, int32_t enable, int32_t reset, double* ms, int64_t* counts);
/* Blocking of the reduction and the back-transformation on this handle:
 * nb_panel reflectors per reduction panel (16, 32, 48 or 64) and nb_back
 * reflectors per back-transformation block (a multiple of 16 in [16, 128]).
 * The defaults (32, 128) are the measured best at n = 4764; the others are
 * supported so the kernels can be checked at every blocking. */
int pph_debug_blocking(void* h, int32_t nb_panel, int32_t nb_back);
/* Products with D(q) and dD/dq in the Rayleigh quotients and velocities:
 * mode 0 dense (matrix-core GEMMs with the Hermitian-dot epilogue), 1
 * automatic (block-sparse when at most a quarter of the force-constant
 * blocks are nonzero, and on precision 'S' with range A, eigenvectors and the
 * FP64 correction of the vectors, whose velocity forms must be FP64; the
 * default), 2 block-sparse whenever the handle built block lists (N <= 2048
 * or sparse).  nblk (may be NULL) receives the number
 * of lower blocks the products now use, 0 for dense.  No counterpart in the
 * reference, whose velocity products are dense (write_output.F90:415-421). */
int pph_debug_sparse(void* h, int32_t mode, int64_t* nblk);
/* Reduction to the tridiagonal: mode 0 one-stage (blocked LATRD/HETRD), 1
 * two-stage (dense -> band of half-bandwidth `band` on the matrix cores, then
 * bulge chasing), 2 the one-launch reduction of small matrices (one
 * workgroup per matrix, the matrix in LDS; orders up to 96 in FP64 and 128 in
 * FP32, larger orders take the one-stage path), -1 automatic (the small
 * kernel where the order fits it); band 16 or 32 (0 keeps the current one).
 * With h == NULL the setting becomes the default of every handle whose own
 * mode is -1 (also the temporary ones pph_heevx makes).  Both reductions feed
 * the same tridiagonal eigensolver and back-transformation; no counterpart
 * in the reference (pzheevx reduces in one stage, diagonalize.F90:73-96). */
int pph_debug_reduction(void* h, int32_t mode, int32_t band);
/* Batch graphs: with range A and profiling off, pph_run captures a batch of a
 * shape its lane has just run eagerly as a HIP graph and replays it for the
 * later batches of that shape (q-points uploaded before each replay, outputs
 * gathered after it).  mode 1 on (the default), 0 off (every batch eager), -1
 * unchanged; captures / replays (may be NULL) receive the counts over the
 * handle's lanes.  No counterpart in the reference. */
int pph_debug_graphs(void* h, int32_t mode, int64_t* captures, int64_t* replays);
/* Tile shape of the matrix-core GEMMs of the eigensolver (reduction updates,
 * band reduction, back-transformation), the autotuner's tile dimension: 1
 * 64x64 (the default), 2 128x64, 3 64x128, 4 128x128 (FP32 only), 0 the
 * process default.  With h == NULL the setting becomes the process default
 * (also for pph_debug_zgemm and the temporary handles of pph_heevx).
 * Products with the dot epilogue and outputs at most 32 columns wide keep
 * their own tiles.  No counterpart in the reference (PBLAS chooses its own
 * blocking inside pzheevx and pzgemm). */
int pph_debug_gemm_tile(void* h, int32_t tile);
/* The FP32 path's Rayleigh-Ritz step over clusters of close eigenvalues
 * (range A): 1 on (the default), 0 off.  Runs of consecutive modes whose FP64
 * Rayleigh quotients lie within 1e-5 max|lambda| (at most 8 modes a cluster,
 * 256 mode pairs a matrix) get the FP64 Ritz values, vectors and velocity
 * forms of their FP32 vectors' span.  With h == NULL the setting becomes the
 * process default (handles created later, the temporary handles of
 * pph_heevx).  No counterpart in the reference
 * (pzheevx computes in FP64 throughout, where such pairs do not mix). */
int pph_debug_cluster_rr(void* h, int32_t mode);
/* FP64 residual-correction steps of the FP32 path's eigenvectors (range A,
 * with eigenvectors): 0, 1 or 2 (the default).  A step computes the
 * residual R = D X - X Theta in FP64 (block-sparse or dense D), M = X^H R and
 * X <- X + X E with E_ij = M_ij / (theta_j - theta_i) for pairs farther apart
 * than the cluster tolerance and E_ii = (1 - |x_i|^2) / 2; the group
 * velocities then come from the corrected vectors.  Each step squares the
 * vectors' error.  The first step's two products are full, on the BF16
 * (bf16 = 1, the default) or FP32 (bf16 = 0) matrix cores; the second works
 * on a band of 16 modes either side of the diagonal, in FP32.  With h == NULL
 * the settings become the process default.  No counterpart in the reference
 * (pzheevx returns FP64 vectors). */
int pph_debug_vector_refinement(void* h, int32_t steps, int32_t bf16);

#ifdef __cplusplus
}
#endif
#endif
