// Real symmetric tridiagonal eigensolver in FP64, batched over matrices:
// Sturm-count bisection for the selected eigenvalues (RANGE A / I / V with
// abstol, as pzheevx's PDSTEBZ stage does) and inverse iteration for the
// eigenvectors (the PZSTEIN stage), with Gram-Schmidt inside clusters of
// close eigenvalues.  Eigenvectors are written interleaved, element i of
// vector k at zt[i * ldzt + k], so neighbouring lanes touch neighbouring
// addresses; that array is the row-major n x M matrix Z_T.
#include "common.hpp"
#include "kernels.hpp"
#include <cfloat>

namespace pph {

constexpr double ULP = DBL_EPSILON;          // 2^-52
constexpr double SAFMIN = DBL_MIN;

__device__ __forceinline__ int64_t sturm_count(const double* __restrict__ d,
                                               const double* __restrict__ e2, int64_t n,
                                               double x, double pivmin) {
  // number of eigenvalues < x (pivots of T - xI below zero)
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int64_t c = (q < 0.0);
  for (int64_t i = 1; i < n; ++i) {
    q = (d[i] - x) - e2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    c += (q < 0.0);
  }
  return c;
}

// Sturm counts at K shifts in one pass over T: K independent recurrences of
// the same form as sturm_count, sharing the loads of d and e^2, so their
// division latencies overlap
template <int K>
__device__ __forceinline__ void sturm_counts(const double* __restrict__ d,
                                             const double* __restrict__ e2, int64_t n,
                                             const double (&x)[K], double pivmin, int64_t (&c)[K]) {
  double q[K];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    q[j] = d[0] - x[j];
    if (fabs(q[j]) < pivmin) q[j] = -pivmin;
    c[j] = (q[j] < 0.0);
  }
  for (int64_t i = 1; i < n; ++i) {
    const double di = d[i], ei = e2[i - 1];
#pragma unroll
    for (int j = 0; j < K; ++j) {
      q[j] = (di - x[j]) - ei / q[j];
      if (fabs(q[j]) < pivmin) q[j] = -pivmin;
      c[j] += (q[j] < 0.0);
    }
  }
}

// one workgroup per matrix
__global__ void __launch_bounds__(256) k_tri_prepare(TriEigWork w) {
  __shared__ double sh[8];
  const int64_t b = blockIdx.x;
  const int64_t n = w.n;
  double* d = w.dd + b * w.strideT;
  double* e = w.ee + b * w.strideT;
  double* e2 = w.e2 + b * w.strideT;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    double di, ei;
    if (w.de_is_float) {
      di = (double)((const float*)w.d)[b * w.strideDE + i];
      ei = (i < n - 1) ? (double)((const float*)w.e)[b * w.strideDE + i] : 0.0;
    } else {
      di = ((const double*)w.d)[b * w.strideDE + i];
      ei = (i < n - 1) ? ((const double*)w.e)[b * w.strideDE + i] : 0.0;
    }
    d[i] = di;
    e[i] = ei;
  }
  __syncthreads();
  double emax = 0.0, gl = 1e300, gu = -1e300;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double ei = e[i];
    const double em = (i > 0) ? e[i - 1] : 0.0;
    const double rad = fabs(em) + fabs(ei);
    gl = fmin(gl, d[i] - rad);
    gu = fmax(gu, d[i] + rad);
    emax = fmax(emax, ei * ei);
  }
  gl = -block_max(-gl, sh);
  gu = block_max(gu, sh);
  emax = block_max(emax, sh);
  const double pivmin = SAFMIN * fmax(1.0, emax);
  const double tnorm = fmax(fabs(gl), fabs(gu));
  // splitting: negligible off-diagonals become exact zeros (PDSTEBZ criterion)
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    double ei = e[i];
    if (i < n - 1) {
      if (fabs(d[i] * d[i + 1]) * ULP * ULP + SAFMIN > ei * ei) ei = 0.0;
    } else {
      ei = 0.0;
    }
    e[i] = ei;
    e2[i] = ei * ei;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double fudge = 2.1;
    gl = gl - fudge * tnorm * ULP * (double)n - fudge * 2.0 * pivmin;
    gu = gu + fudge * tnorm * ULP * (double)n + fudge * pivmin;
    int64_t lo = 0, hi = n;
    if (w.range == 'I') {
      lo = max((int64_t)0, min(n, w.il - 1));
      hi = max(lo, min(n, w.iu));
    } else if (w.range == 'V') {
      lo = (w.vl <= gl) ? 0 : (w.vl >= gu ? n : sturm_count(d, e2, n, w.vl, pivmin));
      hi = (w.vu <= gl) ? 0 : (w.vu >= gu ? n : sturm_count(d, e2, n, w.vu, pivmin));
      // (vl, vu]: eigenvalues equal to vu are counted by the strict count up to ties
      hi = max(lo, hi);
    }
    w.ilo[b] = lo;
    w.ihi[b] = hi;
    double* ti = w.tinfo + b * 4;
    ti[0] = gl; ti[1] = gu; ti[2] = tnorm; ti[3] = pivmin;
    w.info[b] = 0;
  }
}

// one thread per (matrix, selected eigenvalue); d and e^2 are staged in LDS
// when the tridiagonal fits (use_lds), so the Sturm recurrences read LDS.
// Each pass counts at kSects - 1 equally spaced interior points of the
// current interval and keeps the subinterval holding eigenvalue kk, until its
// width is within the tolerance of plain bisection.
constexpr int kSects = 4;
__global__ void __launch_bounds__(256) k_tri_bisect(TriEigWork w, int use_lds) {
  extern __shared__ double tsh[];
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t b = blockIdx.y;
  const int64_t lo_i = w.ilo[b], hi_i = w.ihi[b];
  if ((int64_t)blockIdx.x * blockDim.x >= hi_i - lo_i) return;     // whole block idle
  const double* d = w.dd + b * w.strideT;
  const double* e2 = w.e2 + b * w.strideT;
  if (use_lds) {
    for (int64_t i = threadIdx.x; i < w.n; i += blockDim.x) {
      tsh[i] = d[i];
      tsh[w.n + i] = e2[i];
    }
    __syncthreads();
    d = tsh;
    e2 = tsh + w.n;
  }
  if (k >= hi_i - lo_i) return;
  const int64_t kk = lo_i + k;
  const double* ti = w.tinfo + b * 4;
  const double pivmin = ti[3];
  double lo = ti[0], hi = ti[1];
  const double atol = fmax(fmax(w.abstol, pivmin), w.rtol * ti[2]);
  for (int it = 0; it < 200; ++it) {
    const double tol = fmax(atol, 2.0 * ULP * fmax(fabs(lo), fabs(hi)));
    if (hi - lo <= tol) break;
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    double x[kSects - 1];
    const double h = (hi - lo) / kSects;
#pragma unroll
    for (int j = 0; j < kSects - 1; ++j) x[j] = lo + h * (j + 1);
    if (!(x[0] > lo && x[kSects - 2] < hi)) {
      // too few representable points left: one bisection step
      if (sturm_count(d, e2, w.n, mid, pivmin) > kk) hi = mid;
      else lo = mid;
      continue;
    }
    int64_t c[kSects - 1];
    sturm_counts<kSects - 1>(d, e2, w.n, x, pivmin, c);
    // counts grow with x: the new lower end is the last point counting <= kk
    // eigenvalues below it, the new upper end the first counting more
    double nlo = lo, nhi = hi;
#pragma unroll
    for (int j = 0; j < kSects - 1; ++j)
      if (c[j] <= kk) nlo = x[j];
#pragma unroll
    for (int j = kSects - 2; j >= 0; --j)
      if (c[j] > kk) nhi = x[j];
    if (nlo < nhi) {
      lo = nlo;
      hi = nhi;
    } else if (c[0] > kk) {             // counts out of order (rounding): bisect at x[0]
      hi = x[0];
    } else {
      lo = x[0];
    }
  }
  w.w[b * w.ldw + k] = 0.5 * (lo + hi);
}

__device__ __forceinline__ double rand_pm1(uint64_t s) {
  // splitmix64 -> uniform in (-1, 1)
  s += 0x9E3779B97F4A7C15ull;
  s = (s ^ (s >> 30)) * 0xBF58476D1CE4E5B9ull;
  s = (s ^ (s >> 27)) * 0x94D049BB133111EBull;
  s ^= s >> 31;
  return 2.0 * ((double)(s >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}

// LU with partial pivoting of T - sigma I (the DLAGTF scheme).  Factors are
// stored interleaved (row i of vector k at [i * ld + k]): u0 at U0 = fac,
// u1 at U1 = fac + n ld, the multiplier l at Lm = fac + 2 n ld, and the
// pivot flag as one byte at P = (uint8_t*)(fac + 3 n ld).  The second
// superdiagonal of U is e[i + 1] where row i pivoted and 0 elsewhere, so it is
// read from the tridiagonal instead of being stored.
struct TriFac {
  double* u0;
  double* u1;
  double* l;
  uint8_t* piv;
};
__device__ __forceinline__ TriFac tri_fac(double* fac, int64_t n, int64_t ld) {
  return TriFac{fac, fac + n * ld, fac + 2 * n * ld, reinterpret_cast<uint8_t*>(fac + 3 * n * ld)};
}

__device__ __forceinline__ double u2_of(const double* __restrict__ e, int64_t n, int64_t i, bool p) {
  return (p && i + 2 < n) ? e[i + 1] : 0.0;
}

// Backward substitution with the upper factor (bandwidth 3) in place; returns
// max|y|, and with `norm` also sum y^2 and the entry of largest magnitude.
__device__ __forceinline__ double tri_back(const TriFac& F, const double* __restrict__ e, int64_t n,
                                           int64_t ld, int64_t k, double* __restrict__ y,
                                           bool norm, double& ssq, double& big) {
  double y1 = 0.0, y2 = 0.0, mx = 0.0;
  ssq = 0.0;
  big = 0.0;
  for (int64_t i = n - 1; i >= 0; --i) {
    const double u0 = F.u0[i * ld + k];
    const double u1 = F.u1[i * ld + k];
    const double u2 = u2_of(e, n, i, F.piv[i * ld + k] != 0);
    const double v = (y[i * ld + k] - u1 * y1 - u2 * y2) / u0;
    y[i * ld + k] = v;
    y2 = y1;
    y1 = v;
    mx = fmax(mx, fabs(v));
    if (norm) {
      ssq += v * v;
      if (fabs(v) > fabs(big)) big = v;
    }
  }
  return mx;
}

// Forward elimination (row interchanges and multipliers) of sc * y in place.
__device__ __forceinline__ void tri_forward(const TriFac& F, int64_t n, int64_t ld, int64_t k,
                                            double* __restrict__ y, double sc) {
  double cur = y[k] * sc;
  for (int64_t i = 0; i < n - 1; ++i) {
    const double nxt = y[(i + 1) * ld + k] * sc;
    const double l = F.l[i * ld + k];
    double bi, bn;
    if (F.piv[i * ld + k] == 0) { bi = cur; bn = nxt - l * cur; }
    else { bi = nxt; bn = cur - l * nxt; }
    y[i * ld + k] = bi;
    cur = bn;
  }
  y[(n - 1) * ld + k] = cur;
}

// y <- (T - sigma I)^{-1} y in place; returns max|y|
__device__ double tri_solve(const TriFac& F, const double* __restrict__ e, int64_t n, int64_t ld,
                            int64_t k, double* __restrict__ y) {
  tri_forward(F, n, ld, k, y, 1.0);
  double ssq, big;
  return tri_back(F, e, n, ld, k, y, false, ssq, big);
}

// Factor T - sigma I and, in the same pass, eliminate the random start vector
// (element i = rand_pm1(seed + i)) forward, writing the result to y.
__device__ void tri_factor_start(const double* __restrict__ d, const double* __restrict__ e,
                                 int64_t n, double sigma, double tol, const TriFac& F, int64_t ld,
                                 int64_t k, uint64_t seed, double* __restrict__ y) {
  double alpha = d[0] - sigma, beta = (n > 1) ? e[0] : 0.0;
  double cur = rand_pm1(seed);
  for (int64_t i = 0; i < n - 1; ++i) {
    const double c = e[i];                  // subdiagonal (row i+1, col i)
    const double dn = d[i + 1] - sigma;
    const double en = (i + 2 < n) ? e[i + 1] : 0.0;
    const double nxt = rand_pm1(seed + (uint64_t)(i + 1));
    double u0, u1, l, bi, bn;
    uint8_t piv;
    if (fabs(alpha) >= fabs(c)) {
      u0 = alpha; u1 = beta; piv = 0;
      if (fabs(u0) < tol) u0 = (u0 >= 0.0) ? tol : -tol;
      l = c / u0;
      alpha = dn - l * u1;
      beta = en;
      bi = cur; bn = nxt - l * cur;
    } else {
      u0 = c; u1 = dn; piv = 1;
      l = alpha / c;
      alpha = beta - l * dn;
      beta = -l * en;
      bi = nxt; bn = cur - l * nxt;
    }
    F.u0[i * ld + k] = u0;
    F.u1[i * ld + k] = u1;
    F.l[i * ld + k] = l;
    F.piv[i * ld + k] = piv;
    y[i * ld + k] = bi;
    cur = bn;
  }
  if (fabs(alpha) < tol) alpha = (alpha >= 0.0) ? tol : -tol;
  F.u0[(n - 1) * ld + k] = alpha;
  F.u1[(n - 1) * ld + k] = 0.0;
  F.piv[(n - 1) * ld + k] = 0;
  y[(n - 1) * ld + k] = cur;
}

__device__ double perturbed_shift(const double* __restrict__ wv, int64_t k, double tnorm) {
  // shifts of (nearly) equal eigenvalues are separated by pertol, as DSTEIN does
  const int64_t s = (k > 64) ? k - 64 : 0;
  double x = wv[s];
  for (int64_t i = s + 1; i <= k; ++i) {
    const double pertol = 10.0 * ULP * fmax(fabs(wv[i]), ULP * tnorm);
    x = fmax(wv[i], x + pertol);
  }
  return x;
}

__global__ void __launch_bounds__(256) k_tri_inviter(TriEigWork w, int use_lds) {
  extern __shared__ double tsh[];
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t b = blockIdx.y;
  const int64_t M = w.ihi[b] - w.ilo[b];
  if ((int64_t)blockIdx.x * blockDim.x >= M) return;               // whole block idle
  const int64_t n = w.n, ld = w.ldzt;
  const double* d = w.dd + b * w.strideT;
  const double* e = w.ee + b * w.strideT;
  if (use_lds) {
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
      tsh[i] = d[i];
      tsh[n + i] = e[i];
    }
    __syncthreads();
    d = tsh;
    e = tsh + n;
  }
  if (k >= M) return;
  const double tnorm = w.tinfo[b * 4 + 2];
  const double* wv = w.w + b * w.ldw;
  double* y = w.zt + b * w.strideZt;
  double* fac = w.fac + b * w.strideFac;
  const double sigma = perturbed_shift(wv, k, tnorm);
  const double tol = fmax(ULP * tnorm, SAFMIN);
  const TriFac F = tri_fac(fac, n, ld);
  // w.nsolve inverse-iteration solves from a random start (2 or 3); each
  // solve's scaling by 1 / max|y| is applied as the next forward pass reads y
  const int ns = w.nsolve > 1 ? w.nsolve : 3;
  const uint64_t seed = (uint64_t)(w.ilo[b] + k) * 0x100000001B3ull;
  tri_factor_start(d, e, n, sigma, tol, F, ld, k, seed, y);
  double ssq = 0.0, big = 0.0;
  double mx = tri_back(F, e, n, ld, k, y, false, ssq, big);
  for (int it = 1; it < ns; ++it) {
    tri_forward(F, n, ld, k, y, (mx > 0.0) ? 1.0 / mx : 1.0);
    mx = tri_back(F, e, n, ld, k, y, it == ns - 1, ssq, big);
  }
  // unit norm, largest entry positive
  double sc = 1.0 / sqrt(ssq);
  if (big < 0.0) sc = -sc;
  for (int64_t i = 0; i < n; ++i) y[i * ld + k] *= sc;
}

// one workgroup per cluster of close eigenvalues: redo inverse iteration for
// each member with the earlier members projected out before every solve.
__global__ void __launch_bounds__(256) k_tri_cluster(TriEigWork w, double ctol_rel) {
  __shared__ double sh[8];
  __shared__ double smx;
  const int64_t k0 = blockIdx.x;
  const int64_t b = blockIdx.y;
  const int64_t M = w.ihi[b] - w.ilo[b];
  if (k0 >= M) return;
  const double tnorm = w.tinfo[b * 4 + 2];
  const double ctol = ctol_rel * tnorm;
  const double* wv = w.w + b * w.ldw;
  if (k0 > 0 && wv[k0] - wv[k0 - 1] <= ctol) return;          // not a cluster start
  if (k0 + 1 >= M || wv[k0 + 1] - wv[k0] > ctol) return;       // singleton
  int64_t k1 = k0 + 1;
  while (k1 + 1 < M && wv[k1 + 1] - wv[k1] <= ctol) ++k1;      // members k0..k1
  const int64_t n = w.n, ld = w.ldzt;
  double* y = w.zt + b * w.strideZt;
  const TriFac F = tri_fac(w.fac + b * w.strideFac, n, ld);
  const double* e = w.ee + b * w.strideT;
  for (int64_t k = k0; k <= k1; ++k) {
    const uint64_t seed = (uint64_t)(w.ilo[b] + k) * 0x100000001B3ull + 0x5bd1e995ull;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) y[i * ld + k] = rand_pm1(seed + (uint64_t)i);
    for (int it = 0; it <= 3; ++it) {
      // project out earlier cluster members (modified Gram-Schmidt), twice
      for (int pass = 0; pass < 2; ++pass)
        for (int64_t kp = k0; kp < k; ++kp) {
          double dot = 0.0;
          for (int64_t i = threadIdx.x; i < n; i += blockDim.x) dot += y[i * ld + kp] * y[i * ld + k];
          dot = block_sum(dot, sh);
          for (int64_t i = threadIdx.x; i < n; i += blockDim.x) y[i * ld + k] -= dot * y[i * ld + kp];
          __syncthreads();
        }
      if (it == 3) break;
      __syncthreads();
      if (threadIdx.x == 0) smx = tri_solve(F, e, n, ld, k, y);
      __syncthreads();
      const double sc = (smx > 0.0) ? 1.0 / smx : 1.0;
      for (int64_t i = threadIdx.x; i < n; i += blockDim.x) y[i * ld + k] *= sc;
      __syncthreads();
    }
    double nrm = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) nrm += y[i * ld + k] * y[i * ld + k];
    nrm = block_sum(nrm, sh);
    const double sc = 1.0 / sqrt(nrm);
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) y[i * ld + k] *= sc;
    __syncthreads();
  }
}

void tri_prepare(const TriEigWork& w, hipStream_t s) {
  hipLaunchKernelGGL(k_tri_prepare, dim3((unsigned)w.batch), dim3(256), 0, s, w);
  PPH_CHECK_LAUNCH();
}

// tridiagonals up to this size are staged in LDS (2 n doubles)
constexpr int64_t kTriLdsMax = 9216;

void tri_bisect(const TriEigWork& w, hipStream_t s) {
  const int use_lds = w.n <= kTriLdsMax ? 1 : 0;
  const size_t shm = use_lds ? (size_t)2 * w.n * sizeof(double) : 0;
  dim3 grid((unsigned)ceil_div(w.mmax, 256), (unsigned)w.batch);
  if (use_lds)
    PPH_HIP(hipFuncSetAttribute((const void*)k_tri_bisect,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
  hipLaunchKernelGGL(k_tri_bisect, grid, dim3(256), shm, s, w, use_lds);
  PPH_CHECK_LAUNCH();
}

void tri_vectors(const TriEigWork& w, hipStream_t s) {
  const int use_lds = w.n <= kTriLdsMax ? 1 : 0;
  const size_t shm = use_lds ? (size_t)2 * w.n * sizeof(double) : 0;
  dim3 grid((unsigned)ceil_div(w.mmax, 256), (unsigned)w.batch);
  if (use_lds)
    PPH_HIP(hipFuncSetAttribute((const void*)k_tri_inviter,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
  hipLaunchKernelGGL(k_tri_inviter, grid, dim3(256), shm, s, w, use_lds);
  PPH_CHECK_LAUNCH();
  const double ctol_rel = fmax(w.orfac, 1e-9);
  hipLaunchKernelGGL(k_tri_cluster, dim3((unsigned)w.mmax, (unsigned)w.batch), dim3(256), 0, s, w,
                     ctol_rel);
  PPH_CHECK_LAUNCH();
}

}  // namespace pph
