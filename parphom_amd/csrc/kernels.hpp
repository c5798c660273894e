// Host-side launch interfaces of the HIP kernels.
#pragma once
#include "common.hpp"

#include <algorithm>

namespace pph {

// ---------------------------------------------------------------- assembly (K1, K9)
struct AsmParams {
  const double* phi;        // n x n column-major, symmetrised force constants
  const double* crys;       // N x 3 column-major crystal coordinates
  const double* lat;        // 3 x 3 column-major lattice (row = vector)
  const double* rmass;      // N masses
  const uint16_t* mask;     // N x N minimum-image tie sets
  const double* q;          // nb x 3, row per q-point (reduced coordinates)
  const double* cd_dc;      // 3 x 3 central-difference crystal steps (row per direction)
  double cd_norm;
  int64_t N, nb, mode;      // mode 0 analytic, 1 central difference
  int64_t ldD;              // leading dimension of D64 / D32 (0 = n); dD64 / dD32 use n
  const uint8_t* nzflag;    // N x N, 0 where the block of Phi_sym is zero (null: none known)
  cplx<double>* D64;
  cplx<float>* D32;
  cplx<double>* dD64;
  cplx<float>* dD32;
};

void launch_symmetrise_fc(const double* fc_dev, double* phi, int64_t N, hipStream_t s);
void launch_image_table(const double* crys, const double* pos, const double* lat, uint16_t* mask,
                        int64_t N, hipStream_t s);
void launch_assemble(const AsmParams& P, hipStream_t s);
// flags[a + b N] = 1 when block (a, b) of phi (n x n column-major) is nonzero
void launch_block_nz(const double* phi, uint8_t* flags, int64_t N, hipStream_t s);

// Nonzero 3x3 blocks (brow[k], bcol[k]), brow >= bcol, as a real block times a
// phase sum (assemble.hip): phit[k * 9 + 3 alpha + beta] = Phi~ (per handle);
// bph[(q * nblk + k) * 4 + op] = phase sum of D (op 0) and, with nops = 4, of
// dD/dq_{x,y,z} (ops 1..3) for every q of P.
void launch_block_phi(const double* phi, const uint16_t* mask, const double* rmass,
                      const int32_t* brow, const int32_t* bcol, int64_t nblk, int64_t N,
                      double* phit, hipStream_t s);
void launch_sparse_blocks(const AsmParams& P, const int32_t* brow, const int32_t* bcol, int64_t nblk,
                          int nops, cplx<double>* bph, hipStream_t s);

// ---------------------------------------------------------------- block-sparse quadratic forms (sparse.hip)
// Re(z_i^H Op z_i) for the operators op = op0 .. op0 + nops - 1 (nops 1, 3 or
// 4; op 0 = D, 1..3 = dD/dq_{x,y,z}) and modes i < M, from the blocks of rows
// a (rowptr[a] .. rowptr[a + 1]): Op = phase bph[...op] times phit; z_i from the
// mode-contiguous Zt (row r of matrix b at b * strideT + r * ldt).  Added to
// out0[b * out0_b + i] for op 0 and out1[b * out1_b + (op - 1) * out1_op + i]
// for the others.
template <typename Tz>
void launch_sparse_dots(const int64_t* rowptr, const int32_t* bcol, int64_t nblk,
                        const double* phit, const cplx<double>* bph, int op0, int nops,
                        const cplx<Tz>* Zt, int64_t ldt, int64_t strideT, int64_t N, int64_t M,
                        int64_t batch, double* out0, int64_t out0_b, double* out1,
                        int64_t out1_op, int64_t out1_b, double* ws, hipStream_t s);
// Low-mode Rayleigh-Ritz step and second-order correction of the FP32 path
// on the device (refine.hip): lam (B x mcap) refined in place for the modes
// below lam_low (at most 16 per matrix); C1 = X^H D X(:, 0:16), C2 = X^H
// X(:, 0:16) with leading dimension mcap and matrix stride 16 mcap.
void launch_low_rr(double* lam, const double* nrm, const cplx<double>* C1, const cplx<double>* C2,
                   int64_t M, int64_t mcap, double lam_low, int64_t B, hipStream_t s);
// Rayleigh-Ritz over clusters of close eigenvalues of the FP32 path
// (refine.hip).  Clusters: runs of consecutive modes whose FP64 Rayleigh
// quotients differ by at most kClusterTol max|lambda|, at most kClusterMax
// modes each; a matrix holds at most kClusterSlots combination vectors
// (2 per pair of modes of a cluster), which bounds its clusters.
constexpr int kClusterMax = 8;
constexpr int kClusterSlots = 512;
constexpr int kClusterCount = kClusterSlots / 2;
constexpr double kClusterTol = 1e-5;
// per-matrix cluster list: [0] clusters, [1] slots used, then per cluster
// (first mode, size, first slot), then per slot (mode i, 2 mode j + kind)
constexpr int64_t kClusterMeta = 2 + 3 * kClusterCount + 2 * kClusterSlots;
// the slots a matrix of order n uses (its clusters are fewer than n / 2
// pairs in practice): up to kClusterSlots
inline int64_t cluster_slots(int64_t n) {
  return std::min<int64_t>(kClusterSlots, std::max<int64_t>(64, (2 * n + 63) / 64 * 64));
}
// meta (B x kClusterMeta) from lam (B x mcap, modes 0..M-1), at most `slots`
// slots per matrix
void launch_cluster_find(const double* lam, int64_t M, int64_t mcap, int64_t slots, int32_t* meta,
                         int64_t B, hipStream_t s);
// the combination vectors x_i + x_j (kind 0) and x_i + i x_j (kind 1) of every
// slot p < S in FP64, zero for unused slots: vector p of matrix b, row r at
// dst[b * n * S + r * drs + p * dps]; the modes from the vectors (FP32 or
// FP64), element (r, k) of matrix b at src[b * sstride + r * srs + k * sks]
template <typename Ts>
void launch_cluster_gather(const int32_t* meta, const cplx<Ts>* src, int64_t srs, int64_t sks,
                           int64_t sstride, cplx<double>* dst, int64_t drs, int64_t dps, int64_t n,
                           int64_t S, int64_t B, hipStream_t s);
// The k x k problems of the clusters: H and the velocity forms from the
// diagonal forms (num, vacc: raw forms (b * 3 + a) * mcap + k) and the slot
// forms (numc: b * kClusterSlots + p; vc: (b * 3 + a) * kClusterSlots + p),
// G = X^H X from Z64.  Writes the Ritz values to lam, 1 to den, the Ritz
// forms to vacc (when given), the Ritz vectors to Z64 (n x mcap per matrix,
// ld n) and, when given, to the FP32 copy Z32 (same layout).
void launch_cluster_rr(const int32_t* meta, cplx<double>* Z64, cplx<float>* Z32, int64_t n,
                       int64_t mcap, double* lam, double* den, const double* num,
                       const double* numc, double* vacc, const double* vc, int64_t B,
                       hipStream_t s);
// FP64 residual correction of the vectors (vecref.hip).  Rc = R^H
// mode-contiguous in FP32 (row r of matrix b at b * strideR + r * ldr) with R
// = D X - X Theta from the listed blocks; Xt the vectors (FP32 or FP64)
// mode-contiguous.
template <typename Tx>
void launch_sparse_residual(const int64_t* rowptr, const int32_t* bcol, const int64_t* tptr,
                            const int32_t* trow, const int64_t* tblk, int64_t nblk,
                            const double* phit, const cplx<double>* bph, const cplx<Tx>* Xt,
                            int64_t ldt, int64_t strideT, const double* theta, int64_t ldth,
                            int64_t N, int64_t M, cplx<float>* Rc, int64_t ldr, int64_t strideR,
                            int64_t B, hipStream_t s);
// R = Y - X Theta in FP32 (n x M per matrix, column-major, ld n, stride)
void launch_dense_residual(const cplx<double>* Y, const cplx<double>* X, const double* theta,
                           int64_t ldth, int64_t n, int64_t M, int64_t stride, cplx<float>* R,
                           int64_t B, hipStream_t s);
// E from M = X^H R in place (M x M, ld ldm): M_ij / (theta_j - theta_i) for
// pairs farther apart than tol[b], else 0; (1 - nrm2_i) / 2 on the diagonal;
// with band >= 0, 0 for |i - j| <= band
void launch_vecref_coeffs(cplx<float>* Mx, int64_t ldm, int64_t strideM, int64_t M,
                          const double* theta, const double* nrm2, int64_t ldth, const double* tol,
                          int64_t band, int64_t B, hipStream_t s);
// X64 += W, X32 = X64 rounded (count elements per matrix at stride)
void launch_vecref_update(cplx<double>* X64, cplx<float>* X32, const cplx<float>* W, int64_t count,
                          int64_t stride, int64_t B, hipStream_t s);
// the banded steps: kVecRefBand modes on either side of the diagonal.
// Mb[b * smb + j * (2W + 1) + d] holds M_ij (then E_ij), i = j + d - W.
constexpr int kVecRefBand = 16;
void launch_band_gram(const cplx<float>* X, int64_t ldx, int64_t sx, const cplx<float>* R,
                      int64_t rrs, int64_t rjs, int64_t sr, bool rconj, int64_t n, int64_t M,
                      cplx<float>* Mb, int64_t smb, int64_t B, hipStream_t s);
void launch_band_coeffs(cplx<float>* Mb, int64_t smb, int64_t M, const double* theta,
                        const double* nrm2, int64_t ldth, const double* tol, int64_t B, hipStream_t s);
// Wo = X E over the band (n x M per matrix, column-major, ld n, stride sw),
// or Wo += X E with accumulate
void launch_band_apply(const cplx<float>* X, int64_t ldx, int64_t sx, const cplx<float>* Eb,
                       int64_t smb, int64_t n, int64_t M, cplx<float>* Wo, int64_t sw, bool accumulate,
                       int64_t B, hipStream_t s);
// tol[b] = scale max_k |lam(b, k)|
void launch_lam_tol(const double* lam, int64_t M, int64_t ld, double scale, double* tol, int64_t B,
                    hipStream_t s);
// Spring-model force constants (spring.hip): pairs (pi, pj) with separation
// rvec (3 per pair) and kind 0..2 (intralayer shell: radial p[kind],
// tangential p[3 + kind]) or 3 (interlayer: p[6] exp(-(|r| - inter_r0) /
// inter_decay)); noise amplitude amp, seed of the counter-based noise.  fc:
// dense (N, N, 3, 3) C-order, cleared first; self blocks by the sum rule.
struct SpringParams {
  double p[7];
  double amp, inter_r0, inter_decay;
  uint64_t seed;
};
void launch_spring_fc(int64_t N, int64_t npairs, const int64_t* pi, const int64_t* pj,
                      const double* rvec, const int32_t* kind, const SpringParams& sp, double* fc,
                      hipStream_t s);
// bytes of the workspace launch_sparse_dots needs (per-workgroup partials)
int64_t sparse_dots_ws_bytes(int64_t N, int64_t M, int64_t batch, int nops);
// Y = D X for k columns (n x k per matrix, ld n) from the listed blocks: row
// lists (rowptr, bcol) and their transposes (tptr, trow, tblk: for atom a the
// blocks (c, a) with c > a), phases bph (op 0) and real blocks phit
void launch_sparse_spmm(const int64_t* rowptr, const int32_t* bcol, const int64_t* tptr,
                        const int32_t* trow, const int64_t* tblk, int64_t nblk, const double* phit,
                        const cplx<double>* bph, const cplx<double>* X, int64_t strideX,
                        cplx<double>* Y, int64_t strideY, int64_t N, int64_t k, int64_t batch,
                        hipStream_t s);
// Zt(r, i) = Z(r, i), Z column-major n x m (ldz), Zt with rows ldt apart
template <typename T>
void launch_transpose(const cplx<T>* Z, int64_t n, int64_t m, int64_t ldz, int64_t strideZ,
                      cplx<T>* Zt, int64_t ldt, int64_t strideT, int64_t batch, hipStream_t s);

// ---------------------------------------------------------------- complex GEMM (MFMA)
// C[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b], column-major, batched by stride.
// op: 'N' or 'C' (conjugate transpose).
template <typename T>
struct GemmParams {
  char opA, opB;
  int64_t m, n, k;
  cplx<T> alpha, beta;
  const cplx<T>* A; int64_t lda, strideA;
  const cplx<T>* B; int64_t ldb, strideB;
  // optional second product: K range [k1, k) uses (A2, B2) with the same
  // leading dimensions and ops, batch strides strideA2 / strideB2 (k1 must be a
  // multiple of 16)
  const cplx<T>* A2; const cplx<T>* B2; int64_t k1, strideA2, strideB2;
  cplx<T>* C; int64_t ldc, strideC;
  int64_t batch;
  // inner batch: blockIdx.z = (outer * binner + inner) * ksplit + split; operand
  // offsets add inner * sXin to the outer strides (binner 0 or 1 = no inner batch)
  int64_t binner; int64_t sAin, sBin, sCin, sXin, sDotin;
  // split-K: the K range is cut into ksplit pieces; each piece's product goes
  // to its own slice of ws and a second kernel adds the slices in order
  // (C = sum + beta C).  sCsplit: offset of piece k's C (set by launch_gemm)
  int64_t ksplit, sCsplit;
  // lower != 0: tiles lying entirely above the diagonal of C are skipped
  int lower;
  // diagonal-dot epilogue: when dot_out != null, C is not written; instead
  // dot_out[b*dot_stride + col] += sum_rows Re(conj(X[row, col]) * (op(A) op(B))[row, col])
  // with X = dotX (leading dimension lddx, stride strideX).
  double* dot_out; int64_t dot_stride;
  const cplx<T>* dotX; int64_t lddx, strideX;
  // herm_dot != 0 (with the dot epilogue, opA = 'N', X = op(B)): A is Hermitian
  // and only its lower triangle is read.  The product uses A' = strict lower
  // triangle + Re(diag)/2 and the dot is doubled, since
  // x^H A x = 2 Re(x^H A' x); each row tile stops at its own diagonal, which
  // halves the work.
  int herm_dot;
  // tri != 0 (no dot epilogue, one product): op(A) is triangular and only its
  // nonzero part is read: 1 lower with the diagonal (op(A)(r, k) = 0 for
  // k > r), 2 strictly upper (op(A)(r, k) = 0 for k <= r).  With A Hermitian
  // and stored in its lower triangle, tri 1 with opA 'N' plus tri 2 with opA
  // 'C' (beta 1) give A B from the lower triangle alone.
  int tri;
  // four_m != 0: four real products per complex step instead of three (no
  // cancellation in the imaginary part; a third more matrix-core work)
  int four_m;
  // tile shape of a product without the dot epilogue (kGemmTile*): 0 takes
  // the calling thread's current shape (gemm_tile_scope), which defaults to
  // 64 x 64
  int tile;
  // device workspace of gemm_ws_bytes(p) bytes when that is nonzero: the dot
  // epilogue's per-row-tile partials or split-K's per-piece products, summed
  // in a fixed order so the results do not depend on scheduling
  void* ws;
};

template <typename T> int64_t gemm_ws_bytes(const GemmParams<T>& p);

template <typename T> void launch_gemm(const GemmParams<T>& p, hipStream_t s);
// C = op(A) op(B) on the BF16 matrix cores (gemm_bf16.hip): complex64 in and
// out, inputs rounded to BF16, FP32 accumulation; opA / opB 'N' or 'C', alpha
// 1, beta 0, one plain product (no epilogues, split-K or inner batches)
void launch_cgemm_bf16(const GemmParams<float>& g, hipStream_t s);
// GEMM tile shapes (rows x columns of C per 256-thread workgroup); products
// with the dot epilogue and outputs at most 32 columns wide keep their own
// (64 x 64 and 64 x 32)
enum { kGemmTileDefault = 0, kGemmTile64x64 = 1, kGemmTile128x64 = 2, kGemmTile64x128 = 3,
       kGemmTile128x128 = 4, kGemmTileCount = 5 };
// The shape GEMMs with tile 0 take on the calling host thread while the scope
// lives (each lane runs its batches on its own thread); restored afterwards.
struct gemm_tile_scope {
  explicit gemm_tile_scope(int tile);
  ~gemm_tile_scope();
  int saved;
};
int gemm_current_tile();
// the shape of threads without a scope (initially 64 x 64)
void set_gemm_tile_default(int tile);

// ---------------------------------------------------------------- tridiagonalisation
template <typename T>
struct TridiagWork {
  int64_t n, lda, batch, nb;          // nb = panel width
  cplx<T>* A;  int64_t strideA;       // matrices, overwritten
  cplx<T>* V;  int64_t strideV;       // n x n: all Householder vectors (explicit, unit at j+1)
  cplx<T>* W;  int64_t strideW;       // n x nb panel W
  cplx<T>* tau; int64_t strideTau;    // n
  T* d; T* e; int64_t strideDE;       // diagonal / off-diagonal (n each)
  // per-workgroup partials of the hemv, summed in a fixed order by k_fin:
  // ypart[b * strideY + slot * ldY + l] (l = local row of the trailing matrix),
  // slots 0 .. ntmax - 1 row sums of segment ys, ntmax .. 2 ntmax - 1 column
  // sums of tile row I; dpart[b * strideDp + I * 2 nb + k] the correction dots
  // W^H x, V^H x of tile row I
  cplx<T>* ypart; int64_t strideY, ldY, ntmax;
  cplx<T>* dpart; int64_t strideDp;
  double* scal; int64_t strideScal;   // per-column scalars (tau, scale, alpha2 of the panel)
  double* part; int64_t strideP;      // per-workgroup partials: |x|^2 at [0, n/64 + 1),
  int64_t yvOff;                      //   w^H v (re, im pairs) from part[yvOff]
  int nchunk;                          // hemv column chunks
  hipEvent_t* hemv_ev;                 // optional: event pairs around the hemv of every
};                                     // kHemvTimeStride-th column (j = 0, stride, ...)

constexpr int64_t kHemvTimeStride = 8;

template <typename T> void tridiagonalise(const TridiagWork<T>& w, hipStream_t s);
// hemv tiles of the trailing matrix at column 0 (the most any column has)
inline int64_t hemv_tiles(int64_t n) { return (n - 1 + 15 + 63) / 64; }

// ---------------------------------------------------------------- small matrices (small.hip)
// The whole reduction of a matrix of order n <= kSmallMaxN in one launch, one
// workgroup per matrix with the matrix in LDS.  Uses A (lower triangle read),
// V (vcols columns of n rows per matrix written, zero past column n - 2), tau,
// d, e of w; the panel, partial and scalar buffers are not touched.
template <typename T> constexpr int64_t kSmallMaxN = sizeof(T) == 8 ? 96 : 128;
template <typename T> void tridiagonalise_small(const TridiagWork<T>& w, int64_t vcols, hipStream_t s);

// ---------------------------------------------------------------- two-stage reduction (twostage.hip)
// Dense -> band of half-bandwidth bw (he2hb: panel QR of bw columns, then the
// two-sided update of the trailing matrix on the matrix cores), band -> real
// tridiagonal (hb2st: bulge chasing, one workgroup per matrix with sweeps
// pipelined across its waves).  A must hold a Hermitian matrix in its lower
// triangle; it is destroyed.
template <typename T>
struct TwoStageWork {
  int64_t n, lda, batch, bw;          // bw: 16 or 32
  cplx<T>* A; int64_t strideA;
  // he2hb reflectors: V n x (n + pad) explicit (column c: unit at row c + bw,
  // zeros above), tau n
  cplx<T>* V; int64_t strideV;
  cplx<T>* tau; int64_t strideTau;
  // panel workspaces, leading dimension n: X (= A VT, then Y), VT = V T
  cplx<T>* X; cplx<T>* VT; int64_t strideXV;      // n x bw each
  cplx<T>* Tp; int64_t strideTp;                 // T (bw x bw) then Mx (bw x bw)
  cplx<T>* gpart; int64_t nchunk;                // per-chunk Gram partials: batch x nchunk x bw^2
  // band storage AB[c * 2bw + d] = A(c + d, c), d < 2bw (the bulges reach 2bw - 1)
  cplx<T>* AB; int64_t strideAB;
  // bulge-chasing reflectors: sweep j, step k at V2[(j kmax + k) bw + i], tau2[j kmax + k]
  cplx<T>* V2; cplx<T>* tau2; int64_t kmax, strideV2, strideTau2;
  T* d; T* e; int64_t strideDE;                  // the real tridiagonal
  int* info;                                     // per matrix: bit PPH_INFO_NOCONV on a stalled pipeline
};

template <typename T> void he2hb(const TwoStageWork<T>& w, hipStream_t s);
template <typename T> void hb2st(const TwoStageWork<T>& w, hipStream_t s);
// bulge-chasing steps of sweep j (its reflectors k = 0 .. ntask - 1)
__host__ __device__ inline int64_t hb2st_tasks(int64_t n, int64_t bw, int64_t j) {
  return (n - 2 - j) / bw + 1;
}

// Q2 Z: the bulge-chasing reflectors applied to the columns of Z (n x m,
// leading dimension ldz), in blocks of bw sweeps x one step (compact WY with
// V and VT = V T precomputed per block into Vb and VTb, 2bw x bw each, block
// g kmax + k).
template <typename T>
struct Q2Work {
  int64_t n, batch, bw, kmax, m;
  const cplx<T>* V2; const cplx<T>* tau2; int64_t strideV2, strideTau2;
  cplx<T>* VTb; cplx<T>* Vb; int64_t strideVTb;    // VT and V of every block, same layout
  cplx<T>* Z; int64_t ldz, strideZ;
  int rowmajor;          // Z(row, c) at Z[row ldz + c] instead of Z[row + c ldz]
};
// V and VT of every block (k_q2_prep), then Z <- Q2 Z (k_q2_apply)
template <typename T> void prepare_q2(const Q2Work<T>& w, hipStream_t s);
template <typename T> void apply_q2(const Q2Work<T>& w, hipStream_t s);
template <typename T>
void launch_zt_to_z(const double* zt, int64_t ldzt, int64_t strideZt, cplx<T>* Z, int64_t ldz,
                    int64_t strideZ, int64_t n, int64_t m, int64_t batch, hipStream_t s);
// In place: the row-major FP64 tridiagonal eigenvectors zt (n rows of ldzt)
// become the row-major complex64 Z(i, k) = zt[i ldzt + k] in the same bytes
void launch_zt_to_c64_rows(double* zt, int64_t ldzt, int64_t strideZt, int64_t n, int64_t m,
                           int64_t batch, hipStream_t s);
// Z (column-major, ldz) = Zr (row-major, ldr), complex64
void launch_zr_to_z(const cplx<float>* Zr, int64_t ldr, int64_t strideZr, cplx<float>* Z, int64_t ldz,
                    int64_t strideZ, int64_t n, int64_t m, int64_t batch, hipStream_t s);

// ---------------------------------------------------------------- tridiagonal eigensolver (FP64)
// doubles per matrix for the inverse-iteration factors: u0, u1, l (n x ld
// each) and a pivot byte per element
inline int64_t tri_fac_doubles(int64_t n, int64_t ld) { return 3 * n * ld + (n * ld + 7) / 8; }

struct TriEigWork {
  int64_t n, batch;
  const void* d; const void* e; int64_t strideDE; int de_is_float;
  double* dd; double* ee; double* e2;     // FP64 copies (n each per matrix)
  int64_t strideT;
  char range; double vl, vu; int64_t il, iu; double abstol; double orfac;
  double rtol;                             // bisection stops at rtol * |T| too (0: abstol only)
  int nsolve;                              // inverse-iteration solves per vector (2 or 3)
  int64_t* ilo; int64_t* ihi;              // per matrix index range found [ilo, ihi)
  double* tinfo;                           // per matrix: gl, gu, tnorm, pivmin
  double* w; int64_t ldw;                  // eigenvalues: w[b*ldw + k]
  int want_vectors;
  double* zt; int64_t ldzt; int64_t strideZt;   // row-major n x ldzt eigenvectors of T
  double* fac; int64_t strideFac;               // LU factors: tri_fac_doubles(n, ldzt)
  int64_t* clus; int64_t strideClus;            // cluster starts
  int* info;                                    // per matrix pzheevx-style bits
  int64_t mmax;                                 // max eigenvalues per matrix (threads launched)
};

void tri_prepare(const TriEigWork& w, hipStream_t s);
void tri_bisect(const TriEigWork& w, hipStream_t s);
void tri_vectors(const TriEigWork& w, hipStream_t s);

// ---------------------------------------------------------------- back-transformation
template <typename T>
struct BackWork {
  int64_t n, batch, nb, m;                  // m = eigenvectors to transform
  const cplx<T>* V; int64_t strideV;        // n x n Householder vectors
  const cplx<T>* tau; int64_t strideTau;
  const double* zt; int64_t ldzt; int64_t strideZt;   // row-major T eigenvectors
  cplx<T>* Z; int64_t ldz; int64_t strideZ;            // output n x m column-major
  cplx<T>* Tm; cplx<T>* VT; cplx<T>* Wt; int64_t strideWs;   // workspaces
  // reflector j has its unit at row j + shift (1 after the one-stage
  // reduction, bw after he2hb), so there are n - shift of them
  int64_t shift = 1;
  // Z already holds the vectors to transform (the Z_T conversion skipped)
  bool z_ready = false;
  // workspace of back_transform_ws_bytes bytes (split-K pieces of Wt = V^H Z)
  void* ws = nullptr;
};

// split-K pieces of the product Wt = V^H Z over mr rows for M vectors of B
// matrices: enough workgroups to fill the GPU when the output tiles are few
inline int64_t back_ksplit(int64_t mr, int64_t M, int64_t B) {
  const int64_t tiles = ((M + 63) / 64) * B;
  return std::max<int64_t>(1, std::min<int64_t>((mr + 255) / 256, (1024 + tiles - 1) / tiles));
}
template <typename T>
int64_t back_transform_ws_bytes(int64_t n, int64_t B, int64_t nb, int64_t M, int64_t shift) {
  const int64_t ks = back_ksplit(n - shift, M, B);
  return ks > 1 ? ks * B * nb * M * (int64_t)sizeof(cplx<T>) : 0;
}

template <typename T> void back_transform(const BackWork<T>& w, hipStream_t s);

// ---------------------------------------------------------------- epilogues
void launch_eval_to_cm(const double* lam, double* out, int64_t nb, int64_t m, int64_t ld,
                       hipStream_t s);
void launch_vel_scale(double* v, const double* om_cm, int64_t nb, int64_t m, int64_t ldw,
                      int64_t ldv, hipStream_t s);
template <typename T>
void launch_colnorm2(const cplx<T>* X, int64_t n, int64_t m, int64_t ldx, int64_t strideX,
                     int64_t batch, double* out, int64_t ldo, hipStream_t s);
void launch_rq_finish(const double* num, const double* den, double* lam, int64_t nb, int64_t m,
                      int64_t ld, hipStream_t s);
// Column-wise FP64 helpers (misc.hip).  Element (i, c) of matrix b sits at
// b * stride + c * ld + i, columns c < k; per-column scalars at b * lds + c.
// out = Re sum_i conj(A(i, c)) B(i, c)
void launch_coldot_re(const cplx<double>* A, const cplx<double>* Bm, int64_t n, int64_t k,
                      int64_t ld, int64_t stride, int64_t batch, double* out, int64_t ldo,
                      hipStream_t s);
// Y(:, c) += a sc[c] X(:, c)
void launch_axpy_cols(cplx<double>* Y, const cplx<double>* X, const double* sc, double a, int64_t n,
                      int64_t k, int64_t ld, int64_t stride, int64_t lds, int64_t batch,
                      hipStream_t s);
// T(j, c) *= w[b * ldw + j] for rows j < m (rows of weight 0 are set to 0)
void launch_scale_rows(cplx<double>* T, const double* w, int64_t m, int64_t k, int64_t ld,
                       int64_t stride, int64_t ldw, int64_t batch, hipStream_t s);
// conjugate gradients per column: z += al p, r -= al q with al = rho / pq; then
// p = r + (rho_new / rho_old) p
void launch_cg_update(cplx<double>* z, cplx<double>* r, const cplx<double>* p,
                      const cplx<double>* q, const double* rho, const double* pq, int64_t n,
                      int64_t k, int64_t ld, int64_t stride, int64_t lds, int64_t batch,
                      hipStream_t s);
void launch_cg_direction(cplx<double>* p, const cplx<double>* r, const double* rho_new,
                         const double* rho_old, int64_t n, int64_t k, int64_t ld, int64_t stride,
                         int64_t lds, int64_t batch, hipStream_t s);
template <typename Ti, typename To>
void launch_convert(const cplx<Ti>* in, cplx<To>* out, int64_t count, hipStream_t s);
void launch_fill_zero(void* p, int64_t bytes, hipStream_t s);
// dst[b * dst_ld + i] = src[off[b] + i], i < cnt[b] <= max_cnt (doubles; off,
// cnt device arrays of B entries)
void launch_gather_rows(const double* src, const int64_t* off, const int64_t* cnt, double* dst,
                        int64_t dst_ld, int64_t max_cnt, int64_t B, hipStream_t s);
// info[b] |= flags[b], b < B (device arrays)
void launch_or_flags(int32_t* info, const int* flags, int64_t B, hipStream_t s);

// ---------------------------------------------------------------- density of states (dos.hip)
// freq, func: nk x nb row-major (freq[k * nb + b]; func null: 1 everywhere, the
// reference's default); tri: ntri x 3 row-major, 0-based; E ascending; step: nE
// scratch.  All device pointers.
void launch_dos_lt(const double* freq, const double* func, int64_t nb, const int64_t* tri,
                   int64_t ntri, const double* E, int64_t nE, double* dos, double* nd, double* step,
                   hipStream_t s);
// acoustic sum rule then (M + M^T)/2, `iterations` times, in place (fcsym.hip)
void launch_fc_symmetrize(double* fc, int64_t N, int64_t iterations, hipStream_t s);
// method 1 Gaussian, 2 Lorentzian; freq: nq x nb row-major, each row ascending
void launch_dos_smear(int method, const double* freq, const double* w, int64_t nq, int64_t nb,
                      double width, int64_t natom, const double* E, int64_t nE, double* dos,
                      hipStream_t s);

}  // namespace pph
