// Hermitian -> real symmetric tridiagonal reduction, T = Q^H A Q, batched.
//
// Blocked Householder reduction on the lower triangle: panels of `nb`
// reflectors are formed column by column (the LATRD scheme: the trailing
// matrix is not touched inside a panel; each new column is corrected with the
// panel's V and W), then the trailing matrix gets one rank-2nb update
// A -= V W^H + W V^H on the matrix cores.  Per column three kernels run:
//   k_col   : correct column j with the panel so far, |x|^2 partial sums; the
//             last workgroup forms the reflector (beta, tau, 1/(alpha-beta)).
//   k_hemv  : y = A(j+1:, j+1:) v over a 2-D grid of row blocks x column
//             chunks, accumulated with device atomics, plus the dots W^H v,
//             V^H v; writes v into V.
//   k_fin   : y - V (W^H v) - W (V^H v), w = tau y, and the last workgroup
//             forms alpha2 = -tau/2 (w^H v).
// The final w = tau y + alpha2 v is never written separately: every consumer
// reads W_eff(:, k) = W(:, k) + alpha2_k V(:, k), which avoids a fourth pass.
// Scalars accumulate in FP64 for both FP32 and FP64 matrices.
//
// This is the reduction half of what pzheevx does (diagonalize.F90:92-96).
#include "common.hpp"
#include "kernels.hpp"

namespace pph {

// per-matrix scalar slots
constexpr int S_NORM = 0, S_YV = 1, S_TAU = 3, S_SCALE = 5, S_A2 = 8;

template <typename T>
__device__ __forceinline__ cplx<T> weff(const TridiagWork<T>& w, int64_t b, int64_t r, int k,
                                        int64_t p, const double* scal) {
  const cplx<T> wv = w.W[b * w.strideW + r + (int64_t)k * w.n];
  const cplx<T> vv = w.V[b * w.strideV + r + (p + k) * w.n];
  const cplx<T> a2{(T)scal[S_A2 + 2 * k], (T)scal[S_A2 + 2 * k + 1]};
  return cadd(wv, cmul(a2, vv));
}

// ---------------------------------------------------------------- k_col
template <typename T>
__global__ void __launch_bounds__(64) k_col(TridiagWork<T> w, int64_t j, int jj, int64_t p) {
  __shared__ double sh[8];
  __shared__ cplx<T> vj[64], wj[64];     // row j of V and W_eff for k < jj
  const int64_t b = blockIdx.z;
  const int64_t n = w.n;
  double* scal = w.scal + b * w.strideScal;
  cplx<T>* A = w.A + b * w.strideA;
  const cplx<T>* V = w.V + b * w.strideV;

  if ((int)threadIdx.x < jj) {
    const int k = threadIdx.x;
    vj[k] = V[j + (p + k) * n];
    wj[k] = weff(w, b, j, k, p, scal);
  }
  __syncthreads();

  const int64_t r = j + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  // clear the accumulators k_hemv adds into (y over rows j+1.., dots 2 nb)
  if (r < n) w.ypart[b * w.strideY + r] = cplx<T>{0, 0};
  if (blockIdx.x == 0)
    for (int k = threadIdx.x; k < 2 * (int)w.nb; k += blockDim.x) w.dpart[b * w.strideDp + k] = cplx<T>{0, 0};
  double part = 0.0;
  if (r < n) {
    cplx<T> a = A[r + j * w.lda];
#pragma unroll 4
    for (int k = 0; k < jj; ++k) {
      const cplx<T> vr = V[r + (p + k) * n];
      const cplx<T> wr = weff(w, b, r, k, p, scal);
      // a -= V(r,k) conj(W(j,k)) + W(r,k) conj(V(j,k))
      a = csub(a, cmul(vr, cconj(wj[k])));
      a = csub(a, cmul(wr, cconj(vj[k])));
    }
    if (r == j) a.y = 0;
    A[r + j * w.lda] = a;
    if (r >= j + 2) part = (double)cabs2(a);
  }
  __threadfence();
  part = block_sum(part, sh);
  if (threadIdx.x == 0) {
    atomic_add_relaxed(&scal[S_NORM], part);
    __threadfence();
    const unsigned done = atomic_inc_acqrel(&w.cnt[b * w.strideCnt + 0]);
    if (done == gridDim.x - 1) {
      __threadfence();
      const double xn2 = atomic_load_acq(&scal[S_NORM]);
      const cplx<T> ajj = A[j + j * w.lda];
      w.d[b * w.strideDE + j] = ajj.x;
      double ar = 0, ai = 0;
      if (j + 1 < n) { const cplx<T> al = A[(j + 1) + j * w.lda]; ar = al.x; ai = al.y; }
      double tr = 0, ti = 0, sr = 0, si = 0, beta = ar;
      if (!(xn2 == 0.0 && ai == 0.0)) {
        const double nrm = sqrt(ar * ar + ai * ai + xn2);
        beta = (ar >= 0.0) ? -nrm : nrm;
        tr = (beta - ar) / beta;
        ti = -ai / beta;
        // scale = 1 / (alpha - beta)
        const double dr = ar - beta, di = ai;
        const double den = dr * dr + di * di;
        sr = dr / den;
        si = -di / den;
      }
      w.e[b * w.strideDE + j] = (T)beta;
      w.tau[b * w.strideTau + j] = cplx<T>{(T)tr, (T)ti};
      scal[S_TAU] = tr; scal[S_TAU + 1] = ti;
      scal[S_SCALE] = sr; scal[S_SCALE + 1] = si;
      scal[S_NORM] = 0.0;
      __threadfence();
      atomic_store_rel(&w.cnt[b * w.strideCnt + 0], 0u);
    }
  }
}

// ---------------------------------------------------------------- k_hemv
// grid: x = row blocks of 64 rows over [j+1, n), y = column chunks, z = batch.
template <typename T>
__global__ void __launch_bounds__(256) k_hemv(TridiagWork<T> w, int64_t j, int jj, int64_t p,
                                              int64_t chunk) {
  extern __shared__ char smem_raw[];
  cplx<T>* vs = reinterpret_cast<cplx<T>*>(smem_raw);          // chunk entries
  __shared__ cplx<T> red[4][64];
  const int64_t b = blockIdx.z;
  const int64_t n = w.n, m = n - j - 1;
  const double* scal = w.scal + b * w.strideScal;
  const cplx<T>* A = w.A + b * w.strideA;
  cplx<T>* V = w.V + b * w.strideV;
  const cplx<T> sc{(T)scal[S_SCALE], (T)scal[S_SCALE + 1]};

  const int64_t c0 = blockIdx.y * chunk;                         // offsets within [0, m)
  const int64_t c1 = min(m, c0 + chunk);
  for (int64_t c = c0 + threadIdx.x; c < c1; c += blockDim.x) {
    const int64_t r = j + 1 + c;
    vs[c - c0] = (c == 0) ? cplx<T>{1, 0} : cmul(A[r + j * w.lda], sc);
  }
  __syncthreads();

  const int tr = threadIdx.x % 64, cg = threadIdx.x / 64;
  const int64_t rr = blockIdx.x * 64 + tr;                        // offset within [0, m)
  cplx<T> acc{0, 0};
  if (rr < m) {
    const cplx<T>* Ar = A + (j + 1 + rr) + (j + 1) * w.lda;
    int64_t c = c0 + cg;
#pragma unroll 4
    for (; c < c1; c += 4) {
      const cplx<T> a = Ar[c * w.lda];
      const cplx<T> v = vs[c - c0];
      acc.x += a.x * v.x - a.y * v.y;
      acc.y += a.x * v.y + a.y * v.x;
    }
  }
  red[cg][tr] = acc;
  __syncthreads();
  if (cg == 0 && rr < m) {
    cplx<T> s = red[0][tr];
    s = cadd(s, red[1][tr]);
    s = cadd(s, red[2][tr]);
    s = cadd(s, red[3][tr]);
    cplx<T>* y = w.ypart + b * w.strideY + (j + 1 + rr);
    atomicAdd(&y->x, s.x);
    atomicAdd(&y->y, s.y);
  }
  if (blockIdx.y == 0) {
    // v for this row block; store into V and form correction-dot partials
    cplx<T> vr{0, 0};
    if (rr < m) {
      vr = (rr == 0) ? cplx<T>{1, 0} : cmul(A[(j + 1 + rr) + j * w.lda], sc);
      if (cg == 0) V[(j + 1 + rr) + j * n] = vr;
    }
    // 2*jj dots over 64 rows; wave cg handles k = cg, cg + 4, ...
    for (int k = cg; k < 2 * jj; k += 4) {
      cplx<T> t{0, 0};
      if (rr < m) {
        const cplx<T> x = (k < jj) ? weff(w, b, j + 1 + rr, k, p, scal)
                                   : V[(j + 1 + rr) + (p + k - jj) * n];
        t = cmulc(x, vr);
      }
      t.x = wave_sum(t.x);
      t.y = wave_sum(t.y);
      if (tr == 0) {
        cplx<T>* d = w.dpart + b * w.strideDp + k;
        atomicAdd(&d->x, t.x);
        atomicAdd(&d->y, t.y);
      }
    }
  }
}

// ---------------------------------------------------------------- k_fin
template <typename T>
__global__ void __launch_bounds__(64) k_fin(TridiagWork<T> w, int64_t j, int jj, int64_t p) {
  __shared__ cplx<T> sW[64], sV[64];
  __shared__ double sh[8];
  const int64_t b = blockIdx.z;
  const int64_t n = w.n, m = n - j - 1;
  double* scal = w.scal + b * w.strideScal;
  const cplx<T>* V = w.V + b * w.strideV;
  const cplx<T> tau{(T)scal[S_TAU], (T)scal[S_TAU + 1]};

  for (int k = threadIdx.x; k < 2 * jj; k += blockDim.x) {
    const cplx<T> s = w.dpart[b * w.strideDp + k];
    if (k < jj) sW[k] = s; else sV[k - jj] = s;
  }
  __syncthreads();

  const int64_t rr = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double pr = 0, pi = 0;
  if (rr < m) {
    const int64_t r = j + 1 + rr;
    cplx<T> y = w.ypart[b * w.strideY + r];
#pragma unroll 4
    for (int k = 0; k < jj; ++k) {
      y = csub(y, cmul(V[r + (p + k) * n], sW[k]));
      y = csub(y, cmul(weff(w, b, r, k, p, scal), sV[k]));
    }
    y = cmul(tau, y);
    w.W[b * w.strideW + r + (int64_t)jj * n] = y;
    const cplx<T> v = V[r + j * n];
    const cplx<T> t = cmulc(y, v);
    pr = t.x; pi = t.y;
  }
  __threadfence();
  pr = block_sum(pr, sh);
  pi = block_sum(pi, sh);
  if (threadIdx.x == 0) {
    atomic_add_relaxed(&scal[S_YV], pr);
    atomic_add_relaxed(&scal[S_YV + 1], pi);
    __threadfence();
    const unsigned done = atomic_inc_acqrel(&w.cnt[b * w.strideCnt + 1]);
    if (done == gridDim.x - 1) {
      __threadfence();
      const double yr = atomic_load_acq(&scal[S_YV]), yi = atomic_load_acq(&scal[S_YV + 1]);
      const double tr = scal[S_TAU], ti = scal[S_TAU + 1];
      // alpha2 = -0.5 * tau * (w^H v)
      scal[S_A2 + 2 * jj] = -0.5 * (tr * yr - ti * yi);
      scal[S_A2 + 2 * jj + 1] = -0.5 * (tr * yi + ti * yr);
      scal[S_YV] = 0.0;
      scal[S_YV + 1] = 0.0;
      __threadfence();
      atomic_store_rel(&w.cnt[b * w.strideCnt + 1], 0u);
    }
  }
}

// ---------------------------------------------------------------- panel end
// W(r, k) <- W_eff(r, k) for rows >= r0 and k < kw; W(:, k) = 0 for kw <= k < kpad.
template <typename T>
__global__ void __launch_bounds__(256) k_fixw(TridiagWork<T> w, int64_t r0, int kw, int kpad,
                                              int64_t p) {
  const int64_t b = blockIdx.z;
  const double* scal = w.scal + b * w.strideScal;
  const int64_t r = r0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (r >= w.n || k >= kpad) return;
  cplx<T>* Wc = w.W + b * w.strideW + r + (int64_t)k * w.n;
  if (k < kw) *Wc = weff(w, b, r, k, p, scal);
  else *Wc = cplx<T>{0, 0};
}

template <typename T>
__global__ void k_last_diag(TridiagWork<T> w) {
  const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b >= w.batch) return;
  const int64_t n = w.n;
  w.d[b * w.strideDE + n - 1] = w.A[b * w.strideA + (n - 1) + (n - 1) * w.lda].x;
  w.e[b * w.strideDE + n - 1] = 0;
}

template <typename T>
__global__ void k_reset_scal(TridiagWork<T> w) {
  const int64_t b = blockIdx.x;
  for (int i = threadIdx.x; i < w.strideScal; i += blockDim.x) w.scal[b * w.strideScal + i] = 0.0;
  if (threadIdx.x < 4) w.cnt[b * w.strideCnt + threadIdx.x] = 0u;
}

template <typename T>
void tridiagonalise(const TridiagWork<T>& w, hipStream_t s) {
  const int64_t n = w.n;
  const unsigned B = (unsigned)w.batch;
  if (w.nb > 64 || w.nb % 16 != 0) throw std::runtime_error("tridiagonalise: nb must be 16..64, multiple of 16");
  hipLaunchKernelGGL(k_reset_scal<T>, dim3(B), dim3(256), 0, s, w);
  PPH_CHECK_LAUNCH();
  // V: explicit Householder vectors, zero above the unit entry, n + 64 columns
  launch_fill_zero(w.V, w.batch * w.strideV * (int64_t)sizeof(cplx<T>), s);
  for (int64_t p = 0; p < n - 1; p += w.nb) {
    const int pw = (int)std::min<int64_t>(w.nb, n - 1 - p);
    for (int jj = 0; jj < pw; ++jj) {
      const int64_t j = p + jj;
      const int64_t m = n - j - 1;
      hipLaunchKernelGGL(k_col<T>, dim3((unsigned)ceil_div(n - j, 64), 1, B), dim3(64), 0, s,
                         w, j, jj, p);
      PPH_CHECK_LAUNCH();
      const int64_t nrb = ceil_div(m, 64);
      int64_t nch = std::max<int64_t>(1, std::min<int64_t>((int64_t)w.nchunk,
                                                           ceil_div(2048, nrb * (int64_t)B)));
      nch = std::min<int64_t>(nch, ceil_div(m, 64));
      nch = std::max<int64_t>(nch, ceil_div(m, 2048));        // chunk of v in LDS <= 2048
      const int64_t chunk = ceil_div(m, nch);
      nch = ceil_div(m, chunk);
      const size_t shm = (size_t)chunk * sizeof(cplx<T>);
      hipLaunchKernelGGL(k_hemv<T>, dim3((unsigned)nrb, (unsigned)nch, B), dim3(256), shm, s,
                         w, j, jj, p, chunk);
      PPH_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_fin<T>, dim3((unsigned)ceil_div(m, 64), 1, B), dim3(64), 0, s,
                         w, j, jj, p);
      PPH_CHECK_LAUNCH();
    }
    const int64_t r0 = p + pw;
    const int64_t mt = n - r0;
    if (mt <= 0) continue;
    const int kpad = (int)(ceil_div(pw, 16) * 16);
    hipLaunchKernelGGL(k_fixw<T>, dim3((unsigned)ceil_div(mt, 256), (unsigned)kpad, B), dim3(256),
                       0, s, w, r0, pw, kpad, p);
    PPH_CHECK_LAUNCH();
    // A(r0:, r0:) -= V W^H + W V^H
    GemmParams<T> g{};
    g.opA = 'N'; g.opB = 'C';
    g.m = mt; g.n = mt; g.k = 2 * kpad; g.k1 = kpad;
    g.alpha = cplx<T>{-1, 0}; g.beta = cplx<T>{1, 0};
    g.A = w.V + r0 + p * n; g.lda = n; g.strideA = w.strideV;
    g.B = w.W + r0; g.ldb = n; g.strideB = w.strideW;
    g.A2 = w.W + r0; g.strideA2 = w.strideW;
    g.B2 = w.V + r0 + p * n; g.strideB2 = w.strideV;
    g.C = w.A + r0 + r0 * w.lda; g.ldc = w.lda; g.strideC = w.strideA;
    g.batch = w.batch;
    launch_gemm<T>(g, s);
  }
  hipLaunchKernelGGL(k_last_diag<T>, dim3((unsigned)ceil_div(w.batch, 64)), dim3(64), 0, s, w);
  PPH_CHECK_LAUNCH();
}

template void tridiagonalise<double>(const TridiagWork<double>&, hipStream_t);
template void tridiagonalise<float>(const TridiagWork<float>&, hipStream_t);

}  // namespace pph
