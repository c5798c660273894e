// Low-frequency refinement of the FP32 path on the device.
//
// The frequencies omega = K sqrt(lambda) magnify eigenvalue errors near
// lambda = 0, so after the FP64 Rayleigh quotients the lowest modes of each
// matrix (lambda below the 50 cm^-1 threshold, at most 16) get one
// Rayleigh-Ritz step over the span of all of them and a second-order
// correction from every other mode j outside their cluster,
//   theta_m - sum_j |s_jm|^2 / ((rho_j - theta_m) |x_j|^2),
//   s_jm = sum_c (C1(j, c) - theta_m C2(j, c)) U(c, m),
// with C1 = X^H D X(:, 0:16) and C2 = X^H X(:, 0:16) formed on the matrix
// cores beforehand.  One workgroup per matrix: thread 0 solves the k x k
// generalized problem (Cholesky, then cyclic complex Jacobi), all threads
// form the correction sums, added in a fixed order.  Everything stays on the
// GPU, so a batch needs no host round trip here.
#include "common.hpp"
#include "kernels.hpp"

namespace pph {

constexpr int kLowMax = 16;

// H x = w G x for Hermitian H and positive definite G (k x k, column-major,
// k <= 16), in place in LDS: on return w ascending and U(:, m) the
// G-orthonormal eigenvectors.  G is overwritten by its Cholesky factor.
__device__ void small_gen_eig_dev(int k, cplx<double>* H, cplx<double>* G, cplx<double>* A,
                                  cplx<double>* Q, cplx<double>* U, double* w) {
  auto at = [k](cplx<double>* M, int i, int j) -> cplx<double>& { return M[i + j * k]; };
  // Cholesky G = L L^H (lower)
  for (int j = 0; j < k; ++j) {
    double d = at(G, j, j).x;
    for (int l = 0; l < j; ++l) d -= cabs2(at(G, j, l));
    d = sqrt(fmax(d, 1e-300));
    at(G, j, j) = cplx<double>{d, 0.0};
    for (int i = j + 1; i < k; ++i) {
      cplx<double> s = at(G, i, j);
      for (int l = 0; l < j; ++l) s = csub(s, cmul(at(G, i, l), cconj(at(G, j, l))));
      at(G, i, j) = cplx<double>{s.x / d, s.y / d};
    }
    for (int i = 0; i < j; ++i) at(G, i, j) = cplx<double>{0.0, 0.0};
  }
  // A = L^-1 H L^-H: X = L^-1 H (in U as scratch), then A(:, r) = L^-1 conj(X(r, :))^T
  for (int c = 0; c < k; ++c)
    for (int i = 0; i < k; ++i) {
      cplx<double> s = at(H, i, c);
      for (int l = 0; l < i; ++l) s = csub(s, cmul(at(G, i, l), at(U, l, c)));
      const double d = at(G, i, i).x;
      at(U, i, c) = cplx<double>{s.x / d, s.y / d};
    }
  for (int r = 0; r < k; ++r)
    for (int i = 0; i < k; ++i) {
      cplx<double> s = cconj(at(U, r, i));
      for (int l = 0; l < i; ++l) s = csub(s, cmul(at(G, i, l), at(A, l, r)));
      const double d = at(G, i, i).x;
      at(A, i, r) = cplx<double>{s.x / d, s.y / d};
    }
  for (int i = 0; i < k; ++i) {          // Hermitise
    at(A, i, i).y = 0.0;
    for (int j = i + 1; j < k; ++j) {
      const cplx<double> a = at(A, i, j), b = cconj(at(A, j, i));
      const cplx<double> m{0.5 * (a.x + b.x), 0.5 * (a.y + b.y)};
      at(A, i, j) = m;
      at(A, j, i) = cconj(m);
    }
  }
  for (int i = 0; i < k * k; ++i) Q[i] = cplx<double>{0.0, 0.0};
  for (int i = 0; i < k; ++i) at(Q, i, i) = cplx<double>{1.0, 0.0};
  // Off-diagonal entries below 1e-17 of the diagonal's norm are set to zero
  // instead of rotated away: a rotation computed from an entry at the
  // rounding floor (or from a square that underflows) is not unitary to
  // working precision, and sweeps over such entries let the diagonal drift.
  // A sweep without a rotation ends the iteration.
  for (int sweep = 0; sweep < 30; ++sweep) {
    double tot = 0.0;
    for (int i = 0; i < k; ++i) tot += cabs2(at(A, i, i));
    const double small = 1e-17 * sqrt(tot);
    int rotations = 0;
    for (int p = 0; p < k; ++p)
      for (int q = p + 1; q < k; ++q) {
        const cplx<double> apq = at(A, p, q);
        const double ab = hypot(apq.x, apq.y);
        if (!(ab > small)) {
          at(A, p, q) = cplx<double>{0.0, 0.0};
          at(A, q, p) = cplx<double>{0.0, 0.0};
          continue;
        }
        ++rotations;
        const double app = at(A, p, p).x, aqq = at(A, q, q).x;
        const cplx<double> em{apq.x / ab, -apq.y / ab};     // e^{-i phi}, a_pq = |a_pq| e^{i phi}
        const double tau = (aqq - app) / (2.0 * ab);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        // J = diag(1, e^{-i phi}) [c s; -s c] on the (p, q) plane: A <- J^H A J, Q <- Q J
        for (int r = 0; r < k; ++r) {
          const cplx<double> arp = at(A, r, p), arq = cmul(em, at(A, r, q));
          at(A, r, p) = cplx<double>{c * arp.x - s * arq.x, c * arp.y - s * arq.y};
          at(A, r, q) = cplx<double>{s * arp.x + c * arq.x, s * arp.y + c * arq.y};
        }
        for (int r = 0; r < k; ++r) {
          const cplx<double> apr = at(A, p, r), aqr = cmul(cconj(em), at(A, q, r));
          at(A, p, r) = cplx<double>{c * apr.x - s * aqr.x, c * apr.y - s * aqr.y};
          at(A, q, r) = cplx<double>{s * apr.x + c * aqr.x, s * apr.y + c * aqr.y};
        }
        for (int r = 0; r < k; ++r) {
          const cplx<double> qrp = at(Q, r, p), qrq = cmul(em, at(Q, r, q));
          at(Q, r, p) = cplx<double>{c * qrp.x - s * qrq.x, c * qrp.y - s * qrq.y};
          at(Q, r, q) = cplx<double>{s * qrp.x + c * qrq.x, s * qrp.y + c * qrq.y};
        }
      }
    if (rotations == 0) break;
  }
  // U = L^-H Q, columns in ascending order of the eigenvalues (H reused as scratch)
  for (int c = 0; c < k; ++c)
    for (int i = k - 1; i >= 0; --i) {
      cplx<double> s = at(Q, i, c);
      for (int l = i + 1; l < k; ++l) s = csub(s, cmul(cconj(at(G, l, i)), at(H, l, c)));
      const double d = at(G, i, i).x;
      at(H, i, c) = cplx<double>{s.x / d, s.y / d};
    }
  int ord[kLowMax];
  for (int i = 0; i < k; ++i) ord[i] = i;
  for (int i = 1; i < k; ++i) {          // insertion sort by eigenvalue (stable)
    const int o = ord[i];
    int j = i - 1;
    while (j >= 0 && at(A, ord[j], ord[j]).x > at(A, o, o).x) { ord[j + 1] = ord[j]; --j; }
    ord[j + 1] = o;
  }
  for (int c = 0; c < k; ++c) {
    w[c] = at(A, ord[c], ord[c]).x;
    for (int i = 0; i < k; ++i) at(U, i, c) = at(H, i, ord[c]);
  }
}

// One workgroup per matrix b: lam (B x mcap) in place; nrm = |x_j|^2;
// C1, C2 (B x mcap x 16, column c of matrix b at b * mcap * 16 + c * mcap);
// M modes per matrix, all starting at the bottom of the spectrum.
__global__ void __launch_bounds__(256) k_low_rr(double* __restrict__ lam, const double* __restrict__ nrm,
                                                const cplx<double>* __restrict__ C1,
                                                const cplx<double>* __restrict__ C2, int64_t M,
                                                int64_t mcap, double lam_low) {
  __shared__ cplx<double> H[kLowMax * kLowMax], G[kLowMax * kLowMax], A[kLowMax * kLowMax],
      Q[kLowMax * kLowMax], U[kLowMax * kLowMax];
  __shared__ double th[kLowMax], newl[kLowMax], sh[8];
  __shared__ int kt_s;
  __shared__ int64_t cend_s;
  const int64_t b = blockIdx.x;
  const int tid = threadIdx.x;
  double* lb = lam + b * mcap;
  const double* nb = nrm + b * mcap;
  const cplx<double>* C1b = C1 + b * mcap * kLowMax;
  const cplx<double>* C2b = C2 + b * mcap * kLowMax;
  if (tid == 0) {
    int k = 0;
    const int kmax = M < kLowMax ? (int)M : kLowMax;
    while (k < kmax && lb[k] < lam_low) ++k;
    kt_s = k;
  }
  __syncthreads();
  const int kt = kt_s;
  if (kt == 0) return;
  // modes closer to the last target than 1e-7 max|lambda| belong to its
  // cluster: they are not refined and stay out of the coupling sums
  double mx = 0.0;
  for (int64_t j = tid; j < M; j += blockDim.x) mx = fmax(mx, fabs(lb[j]));
  mx = block_max(mx, sh);
  const double ctol = 1e-7 * fmax(mx, 1e-300);
  for (int i = tid; i < kt * kt; i += blockDim.x) {
    const int r = i % kt, c = i / kt;
    H[i] = C1b[r + c * mcap];
    G[i] = C2b[r + c * mcap];
  }
  if (tid == 0) {
    int64_t ce = kt;
    while (ce < M && lb[ce] - lb[ce - 1] <= ctol) ++ce;
    cend_s = ce;
  }
  __syncthreads();
  if (tid == 0) small_gen_eig_dev(kt, H, G, A, Q, U, th);
  __syncthreads();
  const int64_t cend = cend_s;
  for (int m = 0; m < kt; ++m) {
    double acc = 0.0;
    for (int64_t j = cend + tid; j < M; j += blockDim.x) {
      cplx<double> s{0.0, 0.0};
      for (int c = 0; c < kt; ++c) {
        const cplx<double> a = C1b[j + c * mcap], g = C2b[j + c * mcap];
        s = cadd(s, cmul(cplx<double>{a.x - th[m] * g.x, a.y - th[m] * g.y}, U[c + m * kt]));
      }
      const double dl = lb[j] - th[m];
      if (fabs(dl) > 0.0) acc -= cabs2(s) / (dl * nb[j]);
    }
    acc = block_sum(acc, sh);
    if (tid == 0) newl[m] = th[m] + acc;
  }
  __syncthreads();
  if (tid == 0) {
    for (int i = 1; i < kt; ++i) {
      const double v = newl[i];
      int j = i - 1;
      while (j >= 0 && newl[j] > v) { newl[j + 1] = newl[j]; --j; }
      newl[j + 1] = v;
    }
    for (int m = 0; m < kt; ++m) lb[m] = newl[m];
  }
}

void launch_low_rr(double* lam, const double* nrm, const cplx<double>* C1, const cplx<double>* C2,
                   int64_t M, int64_t mcap, double lam_low, int64_t B, hipStream_t s) {
  if (B <= 0 || M <= 0) return;
  hipLaunchKernelGGL(k_low_rr, dim3((unsigned)B), dim3(256), 0, s, lam, nrm, C1, C2, M, mcap, lam_low);
  PPH_CHECK_LAUNCH();
}

// ---------------------------------------------------------------- clusters
//
// FP32 eigenvectors of modes closer than the FP32 error of the reduction
// (about 1e-7 max|lambda| here) come out as arbitrary mixtures inside their
// cluster, and the FP64 Rayleigh quotients of such mixtures fall anywhere
// between the true eigenvalues.  Each cluster of consecutive modes with
// quotients at most kClusterTol max|lambda| apart gets one Rayleigh-Ritz step
// in FP64 over the span of its FP32 vectors X: H u = theta G u with
// H = X^H D X, G = X^H X; the Ritz vectors X u (unit norm) replace the
// vectors, theta the quotients, and u^H (X^H dD X) u the velocity forms.  A
// mode outside every cluster then keeps an error below about
// (FP32 error)^2 / (kClusterTol max|lambda|).  The off-diagonal entries of
// X^H Op X come from the Hermitian forms f(u) = Re(u^H Op u) of the
// combinations x_i + x_j and x_i + i x_j:
//   Re(x_i^H Op x_j) = (f(x_i + x_j) - f(x_i) - f(x_j)) / 2,
//   Im(x_i^H Op x_j) = (f(x_i) + f(x_j) - f(x_i + i x_j)) / 2,
// so the same form kernels as the Rayleigh numerators (block-sparse forms or
// the GEMM's Hermitian-dot epilogue) evaluate them, in the same FP64
// arithmetic as the diagonal forms they are combined with.

// One workgroup per matrix: thread 0 walks the quotients in order.  A run
// longer than kClusterMax is cut into pieces; a cluster that would overflow
// the slots is skipped (its modes keep their quotients).
__global__ void __launch_bounds__(256) k_cluster_find(const double* __restrict__ lam, int64_t M,
                                                      int64_t mcap, int slots,
                                                      int32_t* __restrict__ meta) {
  __shared__ double sh[8];
  const int64_t b = blockIdx.x;
  const double* lb = lam + b * mcap;
  int32_t* mb = meta + b * kClusterMeta;
  double mx = 0.0;
  for (int64_t j = threadIdx.x; j < M; j += blockDim.x) mx = fmax(mx, fabs(lb[j]));
  mx = block_max(mx, sh);
  if (threadIdx.x != 0) return;
  const double tol = kClusterTol * mx;
  int32_t* cl = mb + 2;
  int32_t* sl = mb + 2 + 3 * kClusterCount;
  int ncl = 0, ns = 0;
  for (int64_t k = 0; k < M;) {
    int64_t e = k + 1;
    while (e < M && e - k < kClusterMax && fabs(lb[e] - lb[e - 1]) <= tol) ++e;
    const int sz = (int)(e - k);
    if (sz > 1 && ns + sz * (sz - 1) <= slots) {
      cl[3 * ncl] = (int32_t)k;
      cl[3 * ncl + 1] = sz;
      cl[3 * ncl + 2] = ns;
      for (int i = 0; i < sz; ++i)
        for (int j = i + 1; j < sz; ++j)
          for (int kind = 0; kind < 2; ++kind) {
            sl[2 * ns] = (int32_t)(k + i);
            sl[2 * ns + 1] = (int32_t)(2 * (k + j) + kind);
            ++ns;
          }
      ++ncl;
    }
    k = e;
  }
  mb[0] = ncl;
  mb[1] = ns;
}

// thread = (row r, slot p), slot_fast: p varies fastest (mode-contiguous output)
template <typename Ts>
__global__ void __launch_bounds__(256) k_cluster_gather(const int32_t* __restrict__ meta,
                                                        const cplx<Ts>* __restrict__ src,
                                                        int64_t srs, int64_t sks, int64_t sstride,
                                                        cplx<double>* __restrict__ dst, int64_t drs,
                                                        int64_t dps, int64_t n, int64_t S,
                                                        int slot_fast) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t b = blockIdx.y;
  if (t >= n * S) return;
  const int64_t r = slot_fast ? t / S : t % n;
  const int64_t p = slot_fast ? t % S : t / n;
  const int32_t* mb = meta + b * kClusterMeta;
  cplx<double> u{0.0, 0.0};
  if (p < mb[1]) {
    const int32_t* sl = mb + 2 + 3 * kClusterCount + 2 * p;
    const int64_t i = sl[0], j = sl[1] >> 1;
    const cplx<Ts>* S = src + b * sstride + r * srs;
    const cplx<Ts> xi = S[i * sks], xj = S[j * sks];
    if (sl[1] & 1) u = cplx<double>{(double)xi.x - (double)xj.y, (double)xi.y + (double)xj.x};
    else u = cplx<double>{(double)xi.x + (double)xj.x, (double)xi.y + (double)xj.y};
  }
  dst[b * n * S + r * drs + p * dps] = u;
}

// index of the entry (i, j), i <= j, in the packed upper triangle of a
// kClusterMax x kClusterMax matrix (row by row)
__device__ __forceinline__ int cl_pidx(int i, int j) {
  return i * (2 * kClusterMax - i + 1) / 2 + (j - i);
}

// grid (cluster, matrix); the workgroups of clusters a matrix does not have
// leave at once
__global__ void __launch_bounds__(256) k_cluster_rr(const int32_t* __restrict__ meta,
                                                    cplx<double>* __restrict__ Z64,
                                                    cplx<float>* __restrict__ Z32, int64_t n,
                                                    int64_t mcap, double* __restrict__ lam,
                                                    double* __restrict__ den,
                                                    const double* __restrict__ num,
                                                    const double* __restrict__ numc,
                                                    double* __restrict__ vacc,
                                                    const double* __restrict__ vc) {
  constexpr int KC = kClusterMax;
  constexpr int NP = KC * (KC + 1) / 2;
  __shared__ cplx<double> H[KC * KC], G[KC * KC], A[KC * KC], Q[KC * KC], U[KC * KC];
  __shared__ cplx<double> Va[3][KC * KC];
  __shared__ double gs[2 * NP], th[KC], sh[8];
  const int c = blockIdx.x;
  const int64_t b = blockIdx.y;
  const int32_t* mb = meta + b * kClusterMeta;
  if (c >= mb[0]) return;
  const int64_t k0 = mb[2 + 3 * c];
  const int s = mb[3 + 3 * c];
  const int q0 = mb[4 + 3 * c];
  const int tid = threadIdx.x;
  cplx<double>* X = Z64 + b * n * mcap + k0 * n;
  // G = X^H X: the packed upper triangle, every thread over its rows (the
  // entries beyond the cluster's size sum zeros)
  cplx<double> acc[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) acc[p] = cplx<double>{0.0, 0.0};
  for (int64_t r = tid; r < n; r += blockDim.x) {
    cplx<double> xr[KC];
#pragma unroll
    for (int i = 0; i < KC; ++i) xr[i] = i < s ? X[r + i * n] : cplx<double>{0.0, 0.0};
#pragma unroll
    for (int i = 0; i < KC; ++i)
#pragma unroll
      for (int j = i; j < KC; ++j) acc[cl_pidx(i, j)] = cadd(acc[cl_pidx(i, j)], cmulc(xr[i], xr[j]));
  }
#pragma unroll
  for (int i = 0; i < KC; ++i)
#pragma unroll
    for (int j = i; j < KC; ++j)
      if (j < s) {                    // s is uniform over the workgroup
        const int p = cl_pidx(i, j);
        const double re = block_sum(acc[p].x, sh);
        const double im = block_sum(acc[p].y, sh);
        if (tid == 0) {
          gs[2 * p] = re;
          gs[2 * p + 1] = im;
        }
      }
  if (tid == 0) {
    const double* nb = num + b * mcap + k0;
    const double* cb = numc + b * kClusterSlots;
    for (int i = 0; i < s; ++i) {
      H[i + i * s] = cplx<double>{nb[i], 0.0};
      G[i + i * s] = cplx<double>{gs[2 * cl_pidx(i, i)], 0.0};
      if (vacc)
        for (int a = 0; a < 3; ++a) Va[a][i + i * s] = cplx<double>{vacc[(b * 3 + a) * mcap + k0 + i], 0.0};
    }
    int q = q0;
    for (int i = 0; i < s; ++i)
      for (int j = i + 1; j < s; ++j, q += 2) {
        const cplx<double> h{0.5 * (cb[q] - nb[i] - nb[j]), 0.5 * (nb[i] + nb[j] - cb[q + 1])};
        H[i + j * s] = h;
        H[j + i * s] = cconj(h);
        const int p = cl_pidx(i, j);
        const cplx<double> g{gs[2 * p], gs[2 * p + 1]};
        G[i + j * s] = g;
        G[j + i * s] = cconj(g);
        if (vacc)
          for (int a = 0; a < 3; ++a) {
            const double* vb = vc + (b * 3 + a) * kClusterSlots;
            const double fi = Va[a][i + i * s].x, fj = Va[a][j + j * s].x;
            const cplx<double> v{0.5 * (vb[q] - fi - fj), 0.5 * (fi + fj - vb[q + 1])};
            Va[a][i + j * s] = v;
            Va[a][j + i * s] = cconj(v);
          }
      }
    small_gen_eig_dev(s, H, G, A, Q, U, th);
    for (int m = 0; m < s; ++m) {
      lam[b * mcap + k0 + m] = th[m];
      den[b * mcap + k0 + m] = 1.0;
    }
    if (vacc)
      for (int a = 0; a < 3; ++a)
        for (int m = 0; m < s; ++m) {
          double f = 0.0;                  // u_m^H Va u_m
          for (int i = 0; i < s; ++i) {
            cplx<double> t{0.0, 0.0};      // (Va u_m)_i
            for (int j = 0; j < s; ++j) t = cadd(t, cmul(Va[a][i + j * s], U[j + m * s]));
            f += cmulc(U[i + m * s], t).x;
          }
          vacc[(b * 3 + a) * mcap + k0 + m] = f;
        }
  }
  __syncthreads();
  // the Ritz vectors X U, row by row (each thread reads its rows before it
  // writes them)
  cplx<float>* X32 = Z32 ? Z32 + b * n * mcap + k0 * n : nullptr;
  for (int64_t r = tid; r < n; r += blockDim.x) {
    cplx<double> xr[KC];
#pragma unroll
    for (int i = 0; i < KC; ++i) xr[i] = i < s ? X[r + i * n] : cplx<double>{0.0, 0.0};
#pragma unroll
    for (int m = 0; m < KC; ++m)
      if (m < s) {
        cplx<double> y{0.0, 0.0};
#pragma unroll
        for (int i = 0; i < KC; ++i)
          if (i < s) y = cadd(y, cmul(xr[i], U[i + m * s]));
        X[r + m * n] = y;
        if (X32) X32[r + m * n] = cplx<float>{(float)y.x, (float)y.y};
      }
  }
}

void launch_cluster_find(const double* lam, int64_t M, int64_t mcap, int64_t slots, int32_t* meta,
                         int64_t B, hipStream_t s) {
  if (B <= 0 || M <= 0) return;
  if (M > INT32_MAX / 2) throw std::runtime_error("launch_cluster_find: too many modes");
  if (slots < 2 || slots > kClusterSlots) throw std::runtime_error("launch_cluster_find: slots out of range");
  hipLaunchKernelGGL(k_cluster_find, dim3((unsigned)B), dim3(256), 0, s, lam, M, mcap, (int)slots, meta);
  PPH_CHECK_LAUNCH();
}

template <typename Ts>
void launch_cluster_gather(const int32_t* meta, const cplx<Ts>* src, int64_t srs, int64_t sks,
                           int64_t sstride, cplx<double>* dst, int64_t drs, int64_t dps, int64_t n,
                           int64_t S, int64_t B, hipStream_t s) {
  if (B <= 0 || n <= 0) return;
  if (S < 2 || S > kClusterSlots) throw std::runtime_error("launch_cluster_gather: slots out of range");
  const int slot_fast = dps == 1 ? 1 : 0;
  hipLaunchKernelGGL(k_cluster_gather<Ts>, dim3((unsigned)ceil_div(n * S, 256), (unsigned)B),
                     dim3(256), 0, s, meta, src, srs, sks, sstride, dst, drs, dps, n, S, slot_fast);
  PPH_CHECK_LAUNCH();
}

template void launch_cluster_gather<float>(const int32_t*, const cplx<float>*, int64_t, int64_t,
                                           int64_t, cplx<double>*, int64_t, int64_t, int64_t,
                                           int64_t, int64_t, hipStream_t);
template void launch_cluster_gather<double>(const int32_t*, const cplx<double>*, int64_t, int64_t,
                                            int64_t, cplx<double>*, int64_t, int64_t, int64_t,
                                            int64_t, int64_t, hipStream_t);

void launch_cluster_rr(const int32_t* meta, cplx<double>* Z64, cplx<float>* Z32, int64_t n,
                       int64_t mcap, double* lam, double* den, const double* num,
                       const double* numc, double* vacc, const double* vc, int64_t B,
                       hipStream_t s) {
  if (B <= 0 || n <= 1) return;
  if (vacc && !vc) throw std::runtime_error("launch_cluster_rr: velocity forms without slot forms");
  hipLaunchKernelGGL(k_cluster_rr, dim3((unsigned)kClusterCount, (unsigned)B), dim3(256), 0, s, meta,
                     Z64, Z32, n, mcap, lam, den, num, numc, vacc, vc);
  PPH_CHECK_LAUNCH();
}

}  // namespace pph
