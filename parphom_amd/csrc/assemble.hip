// K1 / K9: force-constant symmetrisation, minimum-image tables and batched
// dynamical-matrix assembly D(q) with analytic or central-difference dD/dq.
//
// Semantics follow Phonon_Calculations/SRC/dynamical_matrix.F90:200-262
// (create_dynamical_ij) and force_read.F90:187-189, 223-252.
#include "common.hpp"
#include "kernels.hpp"

namespace pph {

// Phi_sym(I, J) = 0.5 Phi_mat(I, J) + 0.5 Phi_mat(J, I) with
// Phi_mat(3i+a, 3j+b) = fc[j][i][b][a] (C-order (N, N, 3, 3) as stored in the
// HDF5 dataset; force_read.F90:223-252 maps (I, J) to Fortran coordinate
// (a, b, i, j)).  Output column-major n x n.
__global__ void k_symmetrise_fc(const double* __restrict__ fc, double* __restrict__ phi,
                                int64_t N) {
  const int64_t n = 3 * N;
  const int64_t I = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t J = blockIdx.y;
  if (I >= n) return;
  const int64_t i = I / 3, a = I % 3, j = J / 3, b = J % 3;
  const double pij = fc[((j * N + i) * 3 + b) * 3 + a];   // Phi_mat(I, J)
  const double pji = fc[((i * N + j) * 3 + a) * 3 + b];   // Phi_mat(J, I)
  phi[I + J * n] = 0.5 * pij + 0.5 * pji;
}

// Minimum-image tie sets (dynamical_matrix.F90:223-237): for each pair (i, j),
// bit l of mask[i + j N] is set when image l of atom j lies within
// tol32 of the smallest of the 9 in-plane image distances.  Ops are kept
// unfused so the distances round exactly as the host restatement does.
__global__ void k_image_table(const double* __restrict__ crys, const double* __restrict__ pos,
                              const double* __restrict__ lat, uint16_t* __restrict__ mask,
                              int64_t N) {
#pragma clang fp contract(off)
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i >= N) return;
  // lat is column-major 3x3 (row index = lattice vector)
  double L[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) L[r][c] = lat[r + 3 * c];
  const double cj0 = crys[j], cj1 = crys[j + N], cj2 = crys[j + 2 * N];
  const double pi0 = pos[i], pi1 = pos[i + N], pi2 = pos[i + 2 * N];
  double dist[9];
  double mn = 1e300;
#pragma unroll
  for (int l = 0; l < 9; ++l) {
    const double t1 = (double)(l / 3 - 1), t2 = (double)(l % 3 - 1);
    const double c0 = cj0 + t1, c1 = cj1 + t2, c2 = cj2;
    const double x = c0 * L[0][0] + c1 * L[1][0] + c2 * L[2][0];
    const double y = c0 * L[0][1] + c1 * L[1][1] + c2 * L[2][1];
    const double z = c0 * L[0][2] + c1 * L[1][2] + c2 * L[2][2];
    const double dx = pi0 - x, dy = pi1 - y, dz = pi2 - z;
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    dist[l] = d;
    mn = fmin(mn, d);
  }
  const double thr = mn + (double)1.0000000133514320e-10;   // f32-rounded 1e-10
  uint16_t m = 0;
#pragma unroll
  for (int l = 0; l < 9; ++l)
    if (dist[l] < thr) m |= (uint16_t)(1u << l);
  mask[i + j * N] = m;
}

// x / b correctly rounded, from r = RN(1 / b): q0 = x r is within one ulp of
// x / b, e = x - q0 b is exact with an FMA, and q0 + e r rounds to RN(x / b)
// (Markstein, IBM J. Res. Develop. 34 (1990) 111-119).  Three instructions
// instead of a full division for each of the 144 quotients a thread forms per
// q-point, with the same result bit for bit.
__device__ __forceinline__ double div_rn(double x, double b, double r) {
  const double q0 = x * r;
  const double e = fma(-q0, b, x);
  return fma(e, r, q0);
}

struct PhaseOut {
  double p0r, p0i;          // sum exp(i theta)
  double pdr[3], pdi[3];    // analytic derivative sums
};

// create_dynamical_ij phase sums for one pair and one q (dynamical_matrix.F90:241-255).
__device__ __forceinline__ void pair_phase(const double* q, double ci0, double ci1, double ci2,
                                           double cj0, double cj1, double cj2, uint16_t m,
                                           const double (&L)[3][3], bool deriv, PhaseOut& o) {
#pragma clang fp contract(off)
  const double m2pi = -2.0 * 3.1415927410125732;   // -2 * f32(pi)
  o.p0r = o.p0i = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) o.pdr[a] = o.pdi[a] = 0.0;
#pragma unroll 1   // images are masked; a rolled loop keeps the kernel at 3 waves/SIMD
  for (int l = 0; l < 9; ++l) {
    if (!(m & (1u << l))) continue;
    const double t1 = (double)(l / 3 - 1), t2 = (double)(l % 3 - 1);
    const double d0 = (cj0 + t1) - ci0, d1 = (cj1 + t2) - ci1, d2 = cj2 - ci2;
    const double pre = q[0] * d0 + q[1] * d1 + q[2] * d2;
    const double th = m2pi * pre;
    double s, c;
    sincos(th, &s, &c);
    o.p0r += c;
    o.p0i += s;
    if (deriv) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const double r = d0 * L[0][a] + d1 * L[1][a] + d2 * L[2][a];
        // -exp(i th) * i * r = r sin(th) - i r cos(th)
        o.pdr[a] += s * r;
        o.pdi[a] -= c * r;
      }
    }
  }
}

// Batched D(q) assembly for all q of the batch.  A workgroup owns 256 atoms i
// (rows 3 i0 .. 3 i0 + 767) against one atom j (columns 3j .. 3j + 2).  Per q
// each thread forms the phase sums of one pair (i, j) into LDS; then each
// thread produces rows tid, tid + 256, tid + 512 of the three columns, so that
// consecutive lanes store consecutive rows (contiguous 4 KB per store in
// complex128).  The thread's nine Phi_sym values are read once for all q.
// Outputs (column-major, leading dimension n, matrix stride n*n):
//   D64 / D32   : D(q) in complex128 / complex64 (either may be null)
//   dD64 / dD32 : 3 matrices per q (x, y, z) of dD/dq (null = not computed)
// mode: 0 analytic derivative, 1 central difference (dynamical_matrix.F90:133-149)
struct AsmArgs {
  const double* phi;
  const double* crys;       // N x 3 column-major
  const double* lat;        // 3 x 3 column-major
  const double* rmass;      // N: mass of each atom's type
  const uint16_t* mask;     // N x N
  const double* q;          // nb x 3 (row per q), reduced coordinates
  const double* cd_dc;      // 3 x 3: crystal step for each Cartesian direction (row a)
  double cd_norm;           // |delta|
  int64_t N, nb, mode;
  int64_t qchunk;           // q-points per workgroup along grid z (blockIdx.z * qchunk ...)
  int64_t ldD;              // leading dimension of D64 / D32 (>= n; dD uses n)
  const uint8_t* nzflag;    // N x N: 0 where the block of Phi_sym is zero (may be null)
  cplx<double>* D64;
  cplx<float>* D32;
  cplx<double>* dD64;
  cplx<float>* dD32;
};

template <int MODE>
__global__ void __launch_bounds__(256) k_assemble(AsmArgs A) {
#pragma clang fp contract(off)
  // per pair: phase sum (2), then either the analytic derivative sums
  // (dr, di) or the central-difference sums at q + dc/2 (dr, di) and q - dc/2 (er, ei)
  __shared__ double sp0r[256], sp0i[256];
  __shared__ double sdr[3][256], sdi[3][256], ser[MODE == 1 ? 3 : 1][256], sei[MODE == 1 ? 3 : 1][256];
  __shared__ double smult[256], ssmm[256], srmult[256], srsmm[256];
  const int64_t N = A.N, n = 3 * N;
  const int tid = threadIdx.x;
  const int64_t i0 = (int64_t)blockIdx.x * 256;
  const int64_t j = blockIdx.y;
  const int64_t i = i0 + tid;                       // the pair whose phase this thread forms
  const int64_t nrow = 3 * min((int64_t)256, N - i0);
  double L[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) L[r][c] = A.lat[r + 3 * c];
  uint16_t m = 0;
  bool nz = true;           // zero blocks give zero elements whatever the phase
  double ci0 = 0, ci1 = 0, ci2 = 0;
  const double cj0 = A.crys[j], cj1 = A.crys[j + N], cj2 = A.crys[j + 2 * N];
  if (i < N) {
    m = A.mask[i + j * N];
    if (A.nzflag) nz = A.nzflag[i + j * N] != 0;
    smult[tid] = (double)__popc((unsigned)m);
    ssmm[tid] = sqrt(A.rmass[i] * A.rmass[j]);
    srmult[tid] = 1.0 / smult[tid];
    srsmm[tid] = 1.0 / ssmm[tid];
    ci0 = A.crys[i]; ci1 = A.crys[i + N]; ci2 = A.crys[i + 2 * N];
  }
  // Phi_sym(3 i0 + r, 3 j + b) for this thread's rows r = tid + 256 k
  double f[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const int64_t r = tid + 256 * k;
      f[k][b] = (r < nrow) ? A.phi[(3 * i0 + r) + (3 * j + b) * n] : 0.0;
    }

  const bool want_d = (A.dD64 != nullptr) || (A.dD32 != nullptr);
  const size_t mstride = (size_t)n * n;
  const int64_t q0 = (int64_t)blockIdx.z * A.qchunk, q1 = min(A.nb, q0 + A.qchunk);
  for (int64_t iq = q0; iq < q1; ++iq) {
    const double* q = A.q + 3 * iq;
    if (i < N && !nz) {
      sp0r[tid] = sp0i[tid] = 0.0;
      for (int a = 0; a < 3; ++a) {
        sdr[a][tid] = sdi[a][tid] = 0.0;
        ser[MODE == 1 ? a : 0][tid] = sei[MODE == 1 ? a : 0][tid] = 0.0;
      }
    } else if (i < N) {
      PhaseOut po;
      pair_phase(q, ci0, ci1, ci2, cj0, cj1, cj2, m, L, want_d && MODE == 0, po);
      sp0r[tid] = po.p0r;
      sp0i[tid] = po.p0i;
      if (MODE == 1 && want_d) {
        for (int a = 0; a < 3; ++a) {
          double qp[3], qm[3];
          for (int c = 0; c < 3; ++c) {
            qp[c] = q[c] + A.cd_dc[3 * a + c] / 2.0;
            qm[c] = q[c] - A.cd_dc[3 * a + c] / 2.0;
          }
          PhaseOut pp, pm;
          pair_phase(qp, ci0, ci1, ci2, cj0, cj1, cj2, m, L, false, pp);
          pair_phase(qm, ci0, ci1, ci2, cj0, cj1, cj2, m, L, false, pm);
          sdr[a][tid] = pp.p0r; sdi[a][tid] = pp.p0i;
          ser[MODE == 1 ? a : 0][tid] = pm.p0r; sei[MODE == 1 ? a : 0][tid] = pm.p0i;
        }
      } else if (want_d) {
        for (int a = 0; a < 3; ++a) { sdr[a][tid] = po.pdr[a]; sdi[a][tid] = po.pdi[a]; }
      }
    }
    __syncthreads();
    const size_t base = (size_t)iq * mstride;
    const size_t baseD = (size_t)iq * A.ldD * n;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int64_t r = tid + 256 * k;
      if (r >= nrow) continue;
      const int pr = (int)(r / 3);                   // pair within the workgroup
      const double p0r = sp0r[pr], p0i = sp0i[pr], mult = smult[pr], smm = ssmm[pr];
      const double rmult = srmult[pr], rsmm = srsmm[pr];
      // ((Phi p) / multiplicity) / sqrt(m_i m_j), each quotient correctly rounded
      auto scale = [&](double fp) { return div_rn(div_rn(fp, mult, rmult), smm, rsmm); };
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const size_t idx = base + (size_t)(3 * i0 + r) + (size_t)(3 * j + b) * n;
        const size_t idxD = baseD + (size_t)(3 * i0 + r) + (size_t)(3 * j + b) * A.ldD;
        const double fv = f[k][b];
        // dij = Phi * phase / multiplicity / sqrt(m_i m_j)
        const double vr = scale(fv * p0r);
        const double vi = scale(fv * p0i);
        if (A.D64) A.D64[idxD] = cplx<double>{vr, vi};
        if (A.D32) A.D32[idxD] = cplx<float>{(float)vr, (float)vi};
        if (want_d) {
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            double wr, wi;
            if (MODE == 1) {
              // (D(q+) - D(q-)) / |delta| elementwise
              const double pr_ = scale(fv * sdr[d][pr]), pi_ = scale(fv * sdi[d][pr]);
              const double mr = scale(fv * ser[MODE == 1 ? d : 0][pr]);
              const double mi = scale(fv * sei[MODE == 1 ? d : 0][pr]);
              wr = (pr_ - mr) / A.cd_norm;
              wi = (pi_ - mi) / A.cd_norm;
            } else {
              wr = scale(fv * sdr[d][pr]);
              wi = scale(fv * sdi[d][pr]);
            }
            const size_t didx = (size_t)(3 * iq + d) * mstride + (idx - base);
            if (A.dD64) A.dD64[didx] = cplx<double>{wr, wi};
            if (A.dD32) A.dD32[didx] = cplx<float>{(float)wr, (float)wi};
          }
        }
      }
    }
    __syncthreads();                                 // LDS reused by the next q
  }
}

// Nonzero 3x3 blocks (a, b), a >= b, of D(q) and dD/dq: each is a complex
// phase sum times the q-independent real block
//   Phi~_ab = (Phi_sym(a, b) / multiplicity) / sqrt(m_a m_b)
// (dynamical_matrix.F90:241-258: the phase depends on the atom pair and the
// tie set only), so a block is stored as Phi~ once per handle plus, per q, the
// phase sums of D and of dD/dq_{x,y,z} (analytic, or the central difference
// (p(q+) - p(q-)) / |delta|).  The quotients are correctly rounded as in
// k_assemble; Phi~ p differs from k_assemble's ((Phi p) / mult) / sqrt(..)
// only by the rounding of the product order.

// phit[k * 9 + 3 alpha + beta] = Phi~(3a + alpha, 3b + beta) for block k
__global__ void __launch_bounds__(256) k_block_phi(const double* __restrict__ phi,
                                                   const uint16_t* __restrict__ mask,
                                                   const double* __restrict__ rmass,
                                                   const int32_t* __restrict__ brow,
                                                   const int32_t* __restrict__ bcol, int64_t nblk,
                                                   int64_t N, double* __restrict__ phit) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= nblk) return;
  const int64_t n = 3 * N, i = brow[k], j = bcol[k];
  const double mult = (double)__popc((unsigned)mask[i + j * N]);
  const double smm = sqrt(rmass[i] * rmass[j]);
  const double rmult = 1.0 / mult, rsmm = 1.0 / smm;
#pragma unroll
  for (int al = 0; al < 3; ++al)
#pragma unroll
    for (int be = 0; be < 3; ++be)
      phit[k * 9 + 3 * al + be] =
          div_rn(div_rn(phi[(3 * i + al) + (3 * j + be) * n], mult, rmult), smm, rsmm);
}

void launch_block_phi(const double* phi, const uint16_t* mask, const double* rmass,
                      const int32_t* brow, const int32_t* bcol, int64_t nblk, int64_t N,
                      double* phit, hipStream_t s) {
  if (nblk <= 0) return;
  hipLaunchKernelGGL(k_block_phi, dim3((unsigned)ceil_div(nblk, 256)), dim3(256), 0, s, phi, mask,
                     rmass, brow, bcol, nblk, N, phit);
  PPH_CHECK_LAUNCH();
}

// bph[(q * nblk + k) * 4 + op]: phase sum of D (op 0) and, with nops = 4, of
// dD/dq_{x,y,z} (ops 1..3) for block k; one thread per (block, q)
template <int MODE>
__global__ void __launch_bounds__(256) k_sparse_blocks(AsmArgs A, const int32_t* __restrict__ brow,
                                                       const int32_t* __restrict__ bcol,
                                                       int64_t nblk, int nops,
                                                       cplx<double>* __restrict__ bph) {
#pragma clang fp contract(off)
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t iq = blockIdx.y;
  if (k >= nblk) return;
  const int64_t N = A.N;
  const int64_t i = brow[k], j = bcol[k];
  double L[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) L[r][c] = A.lat[r + 3 * c];
  const uint16_t m = A.mask[i + j * N];
  const double ci0 = A.crys[i], ci1 = A.crys[i + N], ci2 = A.crys[i + 2 * N];
  const double cj0 = A.crys[j], cj1 = A.crys[j + N], cj2 = A.crys[j + 2 * N];
  const double* q = A.q + 3 * iq;
  const bool deriv = nops > 1;
  PhaseOut po;
  pair_phase(q, ci0, ci1, ci2, cj0, cj1, cj2, m, L, deriv && MODE == 0, po);
  cplx<double>* out = bph + (iq * nblk + k) * 4;
  out[0] = cplx<double>{po.p0r, po.p0i};
  if (!deriv) return;
  for (int a = 0; a < 3; ++a) {
    if (MODE == 1) {
      double qp[3], qm[3];
      for (int c = 0; c < 3; ++c) {
        qp[c] = q[c] + A.cd_dc[3 * a + c] / 2.0;
        qm[c] = q[c] - A.cd_dc[3 * a + c] / 2.0;
      }
      PhaseOut pp, pm;
      pair_phase(qp, ci0, ci1, ci2, cj0, cj1, cj2, m, L, false, pp);
      pair_phase(qm, ci0, ci1, ci2, cj0, cj1, cj2, m, L, false, pm);
      out[1 + a] = cplx<double>{(pp.p0r - pm.p0r) / A.cd_norm, (pp.p0i - pm.p0i) / A.cd_norm};
    } else {
      out[1 + a] = cplx<double>{po.pdr[a], po.pdi[a]};
    }
  }
}

void launch_sparse_blocks(const AsmParams& P, const int32_t* brow, const int32_t* bcol, int64_t nblk,
                          int nops, cplx<double>* bph, hipStream_t s) {
  if (nblk <= 0 || P.nb <= 0) return;
  AsmArgs A{};
  A.phi = P.phi; A.crys = P.crys; A.lat = P.lat; A.rmass = P.rmass; A.mask = P.mask;
  A.q = P.q; A.cd_dc = P.cd_dc; A.cd_norm = P.cd_norm;
  A.N = P.N; A.nb = P.nb; A.mode = P.mode;
  dim3 grid((unsigned)ceil_div(nblk, 256), (unsigned)P.nb);
  if (P.mode == 1)
    hipLaunchKernelGGL(k_sparse_blocks<1>, grid, dim3(256), 0, s, A, brow, bcol, nblk, nops, bph);
  else
    hipLaunchKernelGGL(k_sparse_blocks<0>, grid, dim3(256), 0, s, A, brow, bcol, nblk, nops, bph);
  PPH_CHECK_LAUNCH();
}

// flags[a + b N] = 1 when the 3x3 block (a, b) of Phi_sym has a nonzero entry
// (Phi_sym is symmetric bit for bit, so the flags are too)
__global__ void k_block_nz(const double* __restrict__ phi, uint8_t* __restrict__ flags, int64_t N) {
  const int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t b = blockIdx.y;
  if (a >= N) return;
  const int64_t n = 3 * N;
  bool nz = false;
#pragma unroll
  for (int be = 0; be < 3; ++be)
#pragma unroll
    for (int al = 0; al < 3; ++al) nz |= phi[(3 * a + al) + (3 * b + be) * n] != 0.0;
  flags[a + b * N] = nz ? 1 : 0;
}

void launch_block_nz(const double* phi, uint8_t* flags, int64_t N, hipStream_t s) {
  dim3 grid((unsigned)ceil_div(N, 256), (unsigned)N);
  hipLaunchKernelGGL(k_block_nz, grid, dim3(256), 0, s, phi, flags, N);
  PPH_CHECK_LAUNCH();
}

void launch_symmetrise_fc(const double* fc_dev, double* phi, int64_t N, hipStream_t s) {
  const int64_t n = 3 * N;
  dim3 grid((unsigned)ceil_div(n, 256), (unsigned)n);
  hipLaunchKernelGGL(k_symmetrise_fc, grid, dim3(256), 0, s, fc_dev, phi, N);
  PPH_CHECK_LAUNCH();
}

void launch_image_table(const double* crys, const double* pos, const double* lat, uint16_t* mask,
                        int64_t N, hipStream_t s) {
  dim3 grid((unsigned)ceil_div(N, 256), (unsigned)N);
  hipLaunchKernelGGL(k_image_table, grid, dim3(256), 0, s, crys, pos, lat, mask, N);
  PPH_CHECK_LAUNCH();
}

void launch_assemble(const AsmParams& P, hipStream_t s) {
  AsmArgs A;
  A.phi = P.phi; A.crys = P.crys; A.lat = P.lat; A.rmass = P.rmass; A.mask = P.mask;
  A.q = P.q; A.cd_dc = P.cd_dc; A.cd_norm = P.cd_norm;
  A.N = P.N; A.nb = P.nb; A.mode = P.mode;
  A.ldD = P.ldD > 0 ? P.ldD : 3 * P.N;
  A.nzflag = P.nzflag;
  A.D64 = P.D64; A.D32 = P.D32; A.dD64 = P.dD64; A.dD32 = P.dD32;
  // one workgroup per (256 atoms i, atom j); small cells also split the batch's
  // q-points over grid z so that the launch has a few thousand workgroups
  const int64_t wg = ceil_div(P.N, 256) * P.N;
  int64_t zs = std::min<int64_t>(std::max<int64_t>(P.nb, 1), std::max<int64_t>(1, 4096 / wg));
  A.qchunk = ceil_div(std::max<int64_t>(P.nb, 1), zs);
  zs = ceil_div(std::max<int64_t>(P.nb, 1), A.qchunk);
  dim3 grid((unsigned)ceil_div(P.N, 256), (unsigned)P.N, (unsigned)zs);
  if (A.mode == 1) hipLaunchKernelGGL(k_assemble<1>, grid, dim3(256), 0, s, A);
  else hipLaunchKernelGGL(k_assemble<0>, grid, dim3(256), 0, s, A);
  PPH_CHECK_LAUNCH();
}

}  // namespace pph
