// FP64 residual correction of the FP32 path's eigenvectors.
//
// The FP32 reduction leaves each vector x_k mixed with its neighbours in the
// spectrum, by about eps32 |D| / gap, which the FP64 Rayleigh quotients
// tolerate (second order) but the group velocities <x|dD/dq|x> do not (first
// order).  One correction step, for X with columns x_k and values theta_k:
//   R = D X - X Theta                      (FP64: the residual),
//   M = X^H R,                              M_ij = x_i^H D x_j - theta_j x_i^H x_j,
//   E_ij = M_ij / (theta_j - theta_i)  (i != j, the pair farther apart than
//          the cluster tolerance; closer pairs were settled by the cluster
//          Rayleigh-Ritz step),   E_ii = (1 - |x_i|^2) / 2,
//   X <- X + X E.
// To first order X + X E is orthonormal and diagonalises D, so the error of
// the new vectors is the square of the old one's, (eps32 |D| / gap)^2.  Only
// R needs FP64: M and X E are corrections, of size eps32 |D| and eps32 |D| /
// gap, so FP32 products on the matrix cores (with R and E rounded to FP32)
// carry them to eps32 of their own size.  R comes from the nonzero 3 x 3
// blocks of D (block-sparse force constants) or from an FP64 product with D.
//
// After one full step the remaining error is second order, (eps32 |D| /
// gap)^2, and only pairs of close modes carry a noticeable part of it, so
// later steps restrict M and E to a band of kVecRefBand modes on either side
// of the diagonal (banded kernels below, a few per cent of a full step's
// products).  A banded first step would leave the first-order mixing with
// the far modes, whose sum still moves the velocities by about 1e-5 of the
// largest one here.
#include "common.hpp"
#include "kernels.hpp"

namespace pph {

// Rc = conj(R)^T, mode-contiguous, in FP32: Rc[b * strideR + r * ldr + i] =
// conj((D x_i)(r) - theta_i x_i(r)) for modes i < M.  Xt: the vectors (FP32
// or FP64, widened to FP64), mode-contiguous (row r of matrix b at b * strideT + r * ldt).  Thread = mode
// i of the chunk blockIdx.y, grid.x = atom row a (varying fastest, so the
// workgroups in flight share one chunk of modes and its rows stay in the
// caches): row a of D is its listed blocks (a, c), c <= a, and the Hermitian
// partners of the blocks (c', a), c' > a (transposed lists).
template <typename Tx>
__global__ void __launch_bounds__(256) k_sparse_residual(
    const int64_t* __restrict__ rowptr, const int32_t* __restrict__ bcol,
    const int64_t* __restrict__ tptr, const int32_t* __restrict__ trow,
    const int64_t* __restrict__ tblk, int64_t nblk, const double* __restrict__ phit,
    const cplx<double>* __restrict__ bph, const cplx<Tx>* __restrict__ Xt, int64_t ldt,
    int64_t strideT, const double* __restrict__ theta, int64_t ldth, int64_t M,
    cplx<float>* __restrict__ Rc, int64_t ldr, int64_t strideR) {
  const int64_t i = blockIdx.y * (int64_t)blockDim.x + threadIdx.x;
  const int64_t a = blockIdx.x, b = blockIdx.z;
  if (i >= M) return;
  const cplx<Tx>* X = Xt + b * strideT + i;
  const cplx<double>* P = bph + b * nblk * 4;
  double yr[3] = {0, 0, 0}, yi[3] = {0, 0, 0};
  auto ld = [&](int64_t r) {
    const cplx<Tx> x = X[r * ldt];
    return cplx<double>{(double)x.x, (double)x.y};
  };
  auto add = [&](const double* F, cplx<double> p, int64_t j, bool tr) {
    cplx<double> v[3];
#pragma unroll
    for (int be = 0; be < 3; ++be) v[be] = ld(3 * j + be);
#pragma unroll
    for (int al = 0; al < 3; ++al) {
      double ur = 0.0, ui = 0.0;
#pragma unroll
      for (int be = 0; be < 3; ++be) {
        const double f = tr ? F[3 * be + al] : F[3 * al + be];
        ur += f * v[be].x;
        ui += f * v[be].y;
      }
      yr[al] += p.x * ur - p.y * ui;
      yi[al] += p.x * ui + p.y * ur;
    }
  };
  const int64_t k1 = rowptr[a + 1];
  for (int64_t k = rowptr[a]; k < k1; ++k) add(phit + k * 9, P[k * 4], bcol[k], false);
  const int64_t t1 = tptr[a + 1];
  for (int64_t t = tptr[a]; t < t1; ++t) {
    const int64_t k = tblk[t];
    const cplx<double> p = P[k * 4];
    add(phit + k * 9, cplx<double>{p.x, -p.y}, trow[t], true);
  }
  const double th = theta[b * ldth + i];
  cplx<float>* out = Rc + b * strideR + i;
#pragma unroll
  for (int al = 0; al < 3; ++al) {
    const cplx<double> x = ld(3 * a + al);
    out[(3 * a + al) * ldr] = cplx<float>{(float)(yr[al] - th * x.x), (float)(-(yi[al] - th * x.y))};
  }
}

// R32 = Y - X Theta in FP32, column-major (element (r, i) of matrix b at
// b * n * ld + i * n + r for all three)
__global__ void __launch_bounds__(256) k_dense_residual(const cplx<double>* __restrict__ Y,
                                                        const cplx<double>* __restrict__ X,
                                                        const double* __restrict__ theta,
                                                        int64_t ldth, int64_t n, int64_t M,
                                                        int64_t stride, cplx<float>* __restrict__ R) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t b = blockIdx.y;
  if (t >= n * M) return;
  const int64_t i = t / n;
  const double th = theta[b * ldth + i];
  const cplx<double> y = Y[b * stride + t], x = X[b * stride + t];
  R[b * stride + t] = cplx<float>{(float)(y.x - th * x.x), (float)(y.y - th * x.y)};
}

// E from M in place (M x M per matrix, column-major, ld ldm): E_ij =
// M_ij / (theta_j - theta_i) for pairs farther apart than tol[b], 0 for closer
// ones, E_ii = (1 - nrm2_i) / 2; with band >= 0 the entries with |i - j| <=
// band (the diagonal included) are 0 instead (a banded part supplies them)
__global__ void __launch_bounds__(256) k_vecref_coeffs(cplx<float>* __restrict__ Mx, int64_t ldm,
                                                       int64_t strideM, int64_t M,
                                                       const double* __restrict__ theta,
                                                       const double* __restrict__ nrm2, int64_t ldth,
                                                       const double* __restrict__ tol, int64_t band) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t b = blockIdx.y;
  if (t >= M * M) return;
  const int64_t i = t % M, j = t / M;
  cplx<float>* e = Mx + b * strideM + j * ldm + i;
  if (band >= 0 && i - j <= band && j - i <= band) {
    *e = cplx<float>{0.0f, 0.0f};
    return;
  }
  if (i == j) {
    *e = cplx<float>{(float)(0.5 * (1.0 - nrm2[b * ldth + i])), 0.0f};
    return;
  }
  const double d = theta[b * ldth + j] - theta[b * ldth + i];
  if (!(fabs(d) > tol[b])) {
    *e = cplx<float>{0.0f, 0.0f};
    return;
  }
  const cplx<float> m = *e;
  *e = cplx<float>{(float)((double)m.x / d), (float)((double)m.y / d)};
}

// X64 += W (FP32), X32 = X64 rounded (n x M per matrix, ld n, stride)
__global__ void __launch_bounds__(256) k_vecref_update(cplx<double>* __restrict__ X64,
                                                       cplx<float>* __restrict__ X32,
                                                       const cplx<float>* __restrict__ W,
                                                       int64_t count, int64_t stride) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t b = blockIdx.y;
  if (t >= count) return;
  const int64_t o = b * stride + t;
  const cplx<float> w = W[o];
  cplx<double> x = X64[o];
  x.x += (double)w.x;
  x.y += (double)w.y;
  X64[o] = x;
  X32[o] = cplx<float>{(float)x.x, (float)x.y};
}

// ---------------------------------------------------------------- band
// Banded M: Mb[b * smb + j * ND + d] = sum_r conj(X(r, i)) R(r, j) for
// i = j + d - W, d < ND = 2W + 1 (0 where i is outside [0, M)).  X column-major
// FP32 (X(r, i) at b * sx + i * ldx + r); R(r, j) = Rb[b * sr + r * rrs + j * rjs],
// conjugated when rconj.  Workgroup = 64 consecutive j and all rows, in
// chunks of 32 rows staged through LDS (the X window of 64 + 2W columns and
// the R tile); thread = (j, a quarter of the d range).
template <int W>
__global__ void __launch_bounds__(256) k_band_gram(const cplx<float>* __restrict__ X, int64_t ldx,
                                                   int64_t sx, const cplx<float>* __restrict__ Rb,
                                                   int64_t rrs, int64_t rjs, int64_t sr, int rconj,
                                                   int64_t n, int64_t M, cplx<float>* __restrict__ Mb,
                                                   int64_t smb) {
  constexpr int TJ = 64, RC = 32, WI = TJ + 2 * W, ND = 2 * W + 1, DQ = (ND + 3) / 4;
  __shared__ cplx<float> xs[RC][WI + 1];
  __shared__ cplx<float> rs[RC][TJ + 1];
  const int tid = threadIdx.x;
  const int64_t j0 = (int64_t)blockIdx.x * TJ, b = blockIdx.y, i0 = j0 - W;
  const int jl = tid % TJ, q = tid / TJ, d0 = q * DQ;
  const cplx<float>* Xb = X + b * sx;
  const cplx<float>* Rbb = Rb + b * sr;
  cplx<float> acc[DQ];
#pragma unroll
  for (int k = 0; k < DQ; ++k) acc[k] = cplx<float>{0.0f, 0.0f};
  for (int64_t r0 = 0; r0 < n; r0 += RC) {
    // the X window: rows fastest (column-major source, coalesced)
    for (int e = tid; e < RC * WI; e += 256) {
      const int rr = e % RC, c = e / RC;
      const int64_t r = r0 + rr, i = i0 + c;
      xs[rr][c] = (r < n && i >= 0 && i < M) ? Xb[i * ldx + r] : cplx<float>{0.0f, 0.0f};
    }
    // the R tile, along the unit stride of its layout
    for (int e = tid; e < RC * TJ; e += 256) {
      const int rr = rjs == 1 ? e / TJ : e % RC, c = rjs == 1 ? e % TJ : e / RC;
      const int64_t r = r0 + rr, j = j0 + c;
      cplx<float> v{0.0f, 0.0f};
      if (r < n && j < M) {
        v = Rbb[r * rrs + j * rjs];
        if (rconj) v.y = -v.y;
      }
      rs[rr][c] = v;
    }
    __syncthreads();
    for (int rr = 0; rr < RC; ++rr) {
      const cplx<float> rv = rs[rr][jl];
#pragma unroll
      for (int k = 0; k < DQ; ++k)
        if (d0 + k < ND) acc[k] = cadd(acc[k], cmulc(xs[rr][jl + d0 + k], rv));
    }
    __syncthreads();
  }
  const int64_t j = j0 + jl;
  if (j >= M) return;
#pragma unroll
  for (int k = 0; k < DQ; ++k) {
    const int d = d0 + k;
    if (d >= ND) continue;
    const int64_t i = j + d - W;
    Mb[b * smb + j * ND + d] = (i >= 0 && i < M) ? acc[k] : cplx<float>{0.0f, 0.0f};
  }
}

// E from the banded M in place: E_ij = M_ij / (theta_j - theta_i) for pairs
// farther apart than tol[b], else 0; (1 - nrm2_j) / 2 on the diagonal
template <int W>
__global__ void __launch_bounds__(256) k_band_coeffs(cplx<float>* __restrict__ Mb, int64_t smb,
                                                     int64_t M, const double* __restrict__ theta,
                                                     const double* __restrict__ nrm2, int64_t ldth,
                                                     const double* __restrict__ tol) {
  constexpr int ND = 2 * W + 1;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t b = blockIdx.y;
  if (t >= M * ND) return;
  const int64_t j = t / ND, d = t % ND, i = j + d - W;
  cplx<float>* e = Mb + b * smb + t;
  if (i == j) {
    *e = cplx<float>{(float)(0.5 * (1.0 - nrm2[b * ldth + j])), 0.0f};
    return;
  }
  if (i < 0 || i >= M) return;                // zero already
  const double dl = theta[b * ldth + j] - theta[b * ldth + i];
  if (!(fabs(dl) > tol[b])) {
    *e = cplx<float>{0.0f, 0.0f};
    return;
  }
  const cplx<float> m = *e;
  *e = cplx<float>{(float)((double)m.x / dl), (float)((double)m.y / dl)};
}

// Wo(r, j) = sum_d X(r, j + d - W) E[j][d] (column-major, ld n, stride sw).
// Workgroup = 64 consecutive j, rows in chunks of 32 staged through LDS;
// thread = (j, rows q, q + 4, ...), its E row in registers; the outputs go
// through LDS so each column's rows are written contiguously.
template <int W>
__global__ void __launch_bounds__(256) k_band_apply(const cplx<float>* __restrict__ X, int64_t ldx,
                                                    int64_t sx, const cplx<float>* __restrict__ Eb,
                                                    int64_t smb, int64_t n, int64_t M,
                                                    cplx<float>* __restrict__ Wo, int64_t sw,
                                                    int accumulate) {
  constexpr int TJ = 64, RC = 32, WI = TJ + 2 * W, ND = 2 * W + 1;
  __shared__ cplx<float> xs[RC][WI + 1];
  __shared__ cplx<float> ws[RC][TJ + 1];
  const int tid = threadIdx.x;
  const int64_t j0 = (int64_t)blockIdx.x * TJ, b = blockIdx.y, i0 = j0 - W;
  const int jl = tid % TJ, q = tid / TJ;
  const int64_t j = j0 + jl;
  const cplx<float>* Xb = X + b * sx;
  cplx<float> e[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d)
    e[d] = j < M ? Eb[b * smb + j * ND + d] : cplx<float>{0.0f, 0.0f};
  for (int64_t r0 = 0; r0 < n; r0 += RC) {
    for (int t = tid; t < RC * WI; t += 256) {
      const int rr = t % RC, c = t / RC;
      const int64_t r = r0 + rr, i = i0 + c;
      xs[rr][c] = (r < n && i >= 0 && i < M) ? Xb[i * ldx + r] : cplx<float>{0.0f, 0.0f};
    }
    __syncthreads();
    for (int rr = q; rr < RC; rr += 4) {
      cplx<float> acc{0.0f, 0.0f};
#pragma unroll
      for (int d = 0; d < ND; ++d) acc = cadd(acc, cmul(xs[rr][jl + d], e[d]));
      ws[rr][jl] = acc;
    }
    __syncthreads();
    for (int t = tid; t < RC * TJ; t += 256) {
      const int rr = t % RC, c = t / RC;
      const int64_t r = r0 + rr, jj = j0 + c;
      if (r < n && jj < M) {
        cplx<float>* o = Wo + b * sw + jj * n + r;
        *o = accumulate ? cadd(*o, ws[rr][c]) : ws[rr][c];
      }
    }
    __syncthreads();
  }
}

// tol[b] = scale max_k |lam[b * ld + k]|, k < M
__global__ void __launch_bounds__(256) k_lam_tol(const double* __restrict__ lam, int64_t M,
                                                 int64_t ld, double scale, double* __restrict__ tol) {
  __shared__ double sh[8];
  const int64_t b = blockIdx.x;
  double mx = 0.0;
  for (int64_t k = threadIdx.x; k < M; k += blockDim.x) mx = fmax(mx, fabs(lam[b * ld + k]));
  mx = block_max(mx, sh);
  if (threadIdx.x == 0) tol[b] = scale * mx;
}

template <typename Tx>
void launch_sparse_residual(const int64_t* rowptr, const int32_t* bcol, const int64_t* tptr,
                            const int32_t* trow, const int64_t* tblk, int64_t nblk,
                            const double* phit, const cplx<double>* bph, const cplx<Tx>* Xt,
                            int64_t ldt, int64_t strideT, const double* theta, int64_t ldth,
                            int64_t N, int64_t M, cplx<float>* Rc, int64_t ldr, int64_t strideR,
                            int64_t B, hipStream_t s) {
  if (N <= 0 || M <= 0 || B <= 0) return;
  hipLaunchKernelGGL(k_sparse_residual<Tx>, dim3((unsigned)N, (unsigned)ceil_div(M, 256), (unsigned)B),
                     dim3(256), 0, s, rowptr, bcol, tptr, trow, tblk, nblk, phit, bph, Xt, ldt,
                     strideT, theta, ldth, M, Rc, ldr, strideR);
  PPH_CHECK_LAUNCH();
}

template void launch_sparse_residual<float>(const int64_t*, const int32_t*, const int64_t*,
                                            const int32_t*, const int64_t*, int64_t, const double*,
                                            const cplx<double>*, const cplx<float>*, int64_t, int64_t,
                                            const double*, int64_t, int64_t, int64_t, cplx<float>*,
                                            int64_t, int64_t, int64_t, hipStream_t);
template void launch_sparse_residual<double>(const int64_t*, const int32_t*, const int64_t*,
                                             const int32_t*, const int64_t*, int64_t, const double*,
                                             const cplx<double>*, const cplx<double>*, int64_t,
                                             int64_t, const double*, int64_t, int64_t, int64_t,
                                             cplx<float>*, int64_t, int64_t, int64_t, hipStream_t);

void launch_dense_residual(const cplx<double>* Y, const cplx<double>* X, const double* theta,
                           int64_t ldth, int64_t n, int64_t M, int64_t stride, cplx<float>* R,
                           int64_t B, hipStream_t s) {
  if (n <= 0 || M <= 0 || B <= 0) return;
  hipLaunchKernelGGL(k_dense_residual, dim3((unsigned)ceil_div(n * M, 256), (unsigned)B), dim3(256),
                     0, s, Y, X, theta, ldth, n, M, stride, R);
  PPH_CHECK_LAUNCH();
}

void launch_vecref_coeffs(cplx<float>* Mx, int64_t ldm, int64_t strideM, int64_t M,
                          const double* theta, const double* nrm2, int64_t ldth, const double* tol,
                          int64_t band, int64_t B, hipStream_t s) {
  if (M <= 0 || B <= 0) return;
  hipLaunchKernelGGL(k_vecref_coeffs, dim3((unsigned)ceil_div(M * M, 256), (unsigned)B), dim3(256), 0,
                     s, Mx, ldm, strideM, M, theta, nrm2, ldth, tol, band);
  PPH_CHECK_LAUNCH();
}

void launch_vecref_update(cplx<double>* X64, cplx<float>* X32, const cplx<float>* W, int64_t count,
                          int64_t stride, int64_t B, hipStream_t s) {
  if (count <= 0 || B <= 0) return;
  hipLaunchKernelGGL(k_vecref_update, dim3((unsigned)ceil_div(count, 256), (unsigned)B), dim3(256), 0,
                     s, X64, X32, W, count, stride);
  PPH_CHECK_LAUNCH();
}

void launch_band_gram(const cplx<float>* X, int64_t ldx, int64_t sx, const cplx<float>* R,
                      int64_t rrs, int64_t rjs, int64_t sr, bool rconj, int64_t n, int64_t M,
                      cplx<float>* Mb, int64_t smb, int64_t B, hipStream_t s) {
  if (n <= 0 || M <= 0 || B <= 0) return;
  if (rjs != 1 && rrs != 1) throw std::runtime_error("launch_band_gram: R needs a unit stride");
  hipLaunchKernelGGL((k_band_gram<kVecRefBand>), dim3((unsigned)ceil_div(M, 64), (unsigned)B), dim3(256),
                     0, s, X, ldx, sx, R, rrs, rjs, sr, rconj ? 1 : 0, n, M, Mb, smb);
  PPH_CHECK_LAUNCH();
}

void launch_band_coeffs(cplx<float>* Mb, int64_t smb, int64_t M, const double* theta,
                        const double* nrm2, int64_t ldth, const double* tol, int64_t B, hipStream_t s) {
  if (M <= 0 || B <= 0) return;
  constexpr int64_t ND = 2 * kVecRefBand + 1;
  hipLaunchKernelGGL((k_band_coeffs<kVecRefBand>), dim3((unsigned)ceil_div(M * ND, 256), (unsigned)B),
                     dim3(256), 0, s, Mb, smb, M, theta, nrm2, ldth, tol);
  PPH_CHECK_LAUNCH();
}

void launch_band_apply(const cplx<float>* X, int64_t ldx, int64_t sx, const cplx<float>* Eb,
                       int64_t smb, int64_t n, int64_t M, cplx<float>* Wo, int64_t sw, bool accumulate,
                       int64_t B, hipStream_t s) {
  if (n <= 0 || M <= 0 || B <= 0) return;
  hipLaunchKernelGGL((k_band_apply<kVecRefBand>), dim3((unsigned)ceil_div(M, 64), (unsigned)B),
                     dim3(256), 0, s, X, ldx, sx, Eb, smb, n, M, Wo, sw, accumulate ? 1 : 0);
  PPH_CHECK_LAUNCH();
}

void launch_lam_tol(const double* lam, int64_t M, int64_t ld, double scale, double* tol, int64_t B,
                    hipStream_t s) {
  if (M <= 0 || B <= 0) return;
  hipLaunchKernelGGL(k_lam_tol, dim3((unsigned)B), dim3(256), 0, s, lam, M, ld, scale, tol);
  PPH_CHECK_LAUNCH();
}

}  // namespace pph
