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
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) {
        if (i == j) tot += cabs2(at(A, i, j));
        else off += cabs2(at(A, i, j));
      }
    if (off <= 1e-32 * fmax(tot, 1e-300)) break;
    for (int p = 0; p < k; ++p)
      for (int q = p + 1; q < k; ++q) {
        const cplx<double> apq = at(A, p, q);
        const double ab = sqrt(cabs2(apq));
        if (ab == 0.0) continue;
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

}  // namespace pph
