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
// The trailing matrix is kept correct in its lower triangle only (the hemv
// reads the lower triangle and uses it twice).
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
// 256 threads = 64 rows x 4 waves; wave s handles the panel terms k = s, s+4, ...
template <typename T>
__global__ void __launch_bounds__(256) k_col(TridiagWork<T> w, int64_t j, int jj, int64_t p) {
  __shared__ double sh[8];
  __shared__ cplx<T> vj[64], wj[64];     // row j of V and W_eff for k < jj
  __shared__ cplx<T> red[4][64];
  const int64_t b = blockIdx.z;
  const int64_t n = w.n;
  double* scal = w.scal + b * w.strideScal;
  cplx<T>* A = w.A + b * w.strideA;
  const cplx<T>* V = w.V + b * w.strideV;
  const int lane = threadIdx.x % 64, sub = threadIdx.x / 64;

  if ((int)threadIdx.x < jj) {
    const int k = threadIdx.x;
    vj[k] = V[j + (p + k) * n];
    wj[k] = weff(w, b, j, k, p, scal);
  }
  __syncthreads();

  const int64_t r = j + blockIdx.x * (int64_t)64 + lane;
  // clear the accumulators k_hemv_sym adds into (y over rows j+1.., dots 2 nb)
  if (sub == 0 && r < n) w.ypart[b * w.strideY + r] = cplx<T>{0, 0};
  if (blockIdx.x == 0)
    for (int k = threadIdx.x; k < 2 * (int)w.nb; k += blockDim.x) w.dpart[b * w.strideDp + k] = cplx<T>{0, 0};
  cplx<T> acc{0, 0};
  if (r < n) {
    for (int k = sub; k < jj; k += 4) {
      // V(r,k) conj(W(j,k)) + W(r,k) conj(V(j,k))
      acc = cadd(acc, cmul(V[r + (p + k) * n], cconj(wj[k])));
      acc = cadd(acc, cmul(weff(w, b, r, k, p, scal), cconj(vj[k])));
    }
  }
  red[sub][lane] = acc;
  __syncthreads();
  double part = 0.0;
  if (sub == 0 && r < n) {
    cplx<T> a = A[r + j * w.lda];
    a = csub(a, cadd(cadd(red[0][lane], red[1][lane]), cadd(red[2][lane], red[3][lane])));
    if (r == j) a.y = 0;
    A[r + j * w.lda] = a;
    if (r >= j + 2) part = (double)cabs2(a);
    // the last workgroup reads A(j,j) and A(j+1,j); everything else is read by
    // the next kernel, so only these two writers publish at device scope
    if (r <= j + 1) __threadfence();
  }
  part = block_sum(part, sh);
  if (threadIdx.x == 0) {
    atomic_add_relaxed(&scal[S_NORM], part);
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

// ---------------------------------------------------------------- k_hemv_sym
// y = S v for the Hermitian trailing matrix S = A(j+1:, j+1:), reading only
// its lower triangle once: each element of a 64 x 64 tile S_IJ (J <= I)
// contributes S_IJ v_J to y_I and S_IJ^H v_I to y_J.  grid: x = tile row I,
// y = segment of `seg` tiles along the row, z = batch.  Lane = row of the tile
// row, wave w = columns 16w..16w+15 of every tile; the next tile's elements
// are loaded while the current one is used.  Row sums stay in registers; the
// 16 column sums of a wave are reduced across lanes by a butterfly that halves
// the number of live columns at each step.  Results go to the accumulator
// with device atomics.  Blocks of segment 0 also form the correction dots
// W^H v, V^H v for their rows and write v into V.
constexpr int kHemvSeg = 12;      // tiles per workgroup along a tile row

template <typename T>
__device__ __forceinline__ T xsel(bool h, T a, T b) { return h ? a : b; }

template <typename T>
__global__ void __launch_bounds__(256) k_hemv_sym(TridiagWork<T> w, int64_t j, int jj, int64_t p,
                                                  int seg) {
  __shared__ cplx<T> vseg[kHemvSeg * 64];
  __shared__ cplx<T> red[4][64];
  const int64_t b = blockIdx.z;
  const int64_t n = w.n, m = n - j - 1;
  const int I = blockIdx.x;
  const int J0 = blockIdx.y * seg;
  if (J0 > I) return;
  const int J1 = min(I + 1, J0 + seg);
  const double* scal = w.scal + b * w.strideScal;
  const cplx<T>* A = w.A + b * w.strideA;
  const cplx<T>* S = A + (j + 1) + (j + 1) * w.lda;
  cplx<T>* V = w.V + b * w.strideV;
  cplx<T>* y = w.ypart + b * w.strideY + (j + 1);
  const cplx<T> sc{(T)scal[S_SCALE], (T)scal[S_SCALE + 1]};
  const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
  auto vval = [&](int64_t o) -> cplx<T> {
    if (o >= m) return cplx<T>{0, 0};
    return (o == 0) ? cplx<T>{1, 0} : cmul(A[(j + 1 + o) + j * w.lda], sc);
  };
  const int64_t rI = (int64_t)I * 64 + lane;          // this lane's row within S
  const cplx<T> vI = vval(rI);
  for (int o = tid; o < (J1 - J0) * 64; o += blockDim.x) vseg[o] = vval((int64_t)J0 * 64 + o);

  cplx<T> cur[16], nxt[16];
  auto load_tile = [&](int J, cplx<T>(&dst)[16]) {
#pragma unroll
    for (int cc = 0; cc < 16; ++cc) {
      const int64_t col = (int64_t)J * 64 + wv * 16 + cc;
      const bool ok = (rI < m) && (col < m);
      const cplx<T> a = S[(ok ? rI : 0) + (ok ? col : 0) * w.lda];
      dst[cc] = ok ? a : cplx<T>{0, 0};
    }
  };
  load_tile(J0, nxt);
  __syncthreads();                                   // vseg ready
  cplx<T> rowacc{0, 0};
  const bool h5 = (lane >> 5) & 1, h4 = (lane >> 4) & 1, h3 = (lane >> 3) & 1, h2 = (lane >> 2) & 1;
  const int mycol = wv * 16 + (h5 ? 8 : 0) + (h4 ? 4 : 0) + (h3 ? 2 : 0) + (h2 ? 1 : 0);
  for (int J = J0; J < J1; ++J) {
#pragma unroll
    for (int cc = 0; cc < 16; ++cc) cur[cc] = nxt[cc];
    if (J + 1 < J1) load_tile(J + 1, nxt);
    const bool diag = (J == I);
    const cplx<T>* vJ = vseg + (J - J0) * 64 + wv * 16;
    T pr[16], pi[16];
#pragma unroll
    for (int cc = 0; cc < 16; ++cc) {
      const int c = wv * 16 + cc;
      const cplx<T> a = cur[cc];
      // row sum: lower triangle including the (real) diagonal
      cplx<T> ar = a;
      if (diag && c == lane) ar.y = 0;
      if (!(diag && c > lane)) rowacc = cadd(rowacc, cmul(ar, vJ[cc]));
      // column sum: strictly lower elements, conj(S(r, c)) v(r)
      const cplx<T> q = (diag && lane <= c) ? cplx<T>{0, 0} : cmulc(a, vI);
      pr[cc] = q.x;
      pi[cc] = q.y;
    }
    // butterfly reduce-scatter of the 16 column sums over the 64 lanes
    T r8[8], i8[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      r8[k] = xsel(h5, pr[8 + k], pr[k]) + __shfl_xor(xsel(h5, pr[k], pr[8 + k]), 32, WAVE);
      i8[k] = xsel(h5, pi[8 + k], pi[k]) + __shfl_xor(xsel(h5, pi[k], pi[8 + k]), 32, WAVE);
    }
    T r4[4], i4[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      r4[k] = xsel(h4, r8[4 + k], r8[k]) + __shfl_xor(xsel(h4, r8[k], r8[4 + k]), 16, WAVE);
      i4[k] = xsel(h4, i8[4 + k], i8[k]) + __shfl_xor(xsel(h4, i8[k], i8[4 + k]), 16, WAVE);
    }
    T r2[2], i2[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      r2[k] = xsel(h3, r4[2 + k], r4[k]) + __shfl_xor(xsel(h3, r4[k], r4[2 + k]), 8, WAVE);
      i2[k] = xsel(h3, i4[2 + k], i4[k]) + __shfl_xor(xsel(h3, i4[k], i4[2 + k]), 8, WAVE);
    }
    T r1 = xsel(h2, r2[1], r2[0]) + __shfl_xor(xsel(h2, r2[0], r2[1]), 4, WAVE);
    T i1 = xsel(h2, i2[1], i2[0]) + __shfl_xor(xsel(h2, i2[0], i2[1]), 4, WAVE);
    r1 += __shfl_xor(r1, 2, WAVE);
    i1 += __shfl_xor(i1, 2, WAVE);
    r1 += __shfl_xor(r1, 1, WAVE);
    i1 += __shfl_xor(i1, 1, WAVE);
    const int64_t cJ = (int64_t)J * 64 + mycol;
    if ((lane & 3) == 0 && cJ < m) {
      atomicAdd(&y[cJ].x, r1);
      atomicAdd(&y[cJ].y, i1);
    }
  }
  red[wv][lane] = rowacc;
  __syncthreads();
  if (wv == 0 && rI < m) {
    const cplx<T> s = cadd(cadd(red[0][lane], red[1][lane]), cadd(red[2][lane], red[3][lane]));
    atomicAdd(&y[rI].x, s.x);
    atomicAdd(&y[rI].y, s.y);
  }
  if (blockIdx.y == 0) {
    // v for this tile row: store into V and form the correction dots
    const cplx<T> vr = (rI < m) ? vI : cplx<T>{0, 0};
    if (wv == 0 && rI < m) V[(j + 1 + rI) + j * n] = vr;
    for (int k = wv; k < 2 * jj; k += 4) {
      cplx<T> t{0, 0};
      if (rI < m) {
        const cplx<T> x = (k < jj) ? weff(w, b, j + 1 + rI, k, p, scal)
                                   : V[(j + 1 + rI) + (p + k - jj) * n];
        t = cmulc(x, vr);
      }
      t.x = wave_sum(t.x);
      t.y = wave_sum(t.y);
      if (lane == 0) {
        cplx<T>* d = w.dpart + b * w.strideDp + k;
        atomicAdd(&d->x, t.x);
        atomicAdd(&d->y, t.y);
      }
    }
  }
}

// ---------------------------------------------------------------- k_fin
// 256 threads = 64 rows x 4 waves; wave s handles the panel terms k = s, s+4, ...
template <typename T>
__global__ void __launch_bounds__(256) k_fin(TridiagWork<T> w, int64_t j, int jj, int64_t p) {
  __shared__ cplx<T> sW[64], sV[64];
  __shared__ cplx<T> red[4][64];
  __shared__ double sh[8];
  const int64_t b = blockIdx.z;
  const int64_t n = w.n, m = n - j - 1;
  double* scal = w.scal + b * w.strideScal;
  const cplx<T>* V = w.V + b * w.strideV;
  const cplx<T> tau{(T)scal[S_TAU], (T)scal[S_TAU + 1]};
  const int lane = threadIdx.x % 64, sub = threadIdx.x / 64;

  for (int k = threadIdx.x; k < 2 * jj; k += blockDim.x) {
    const cplx<T> s = w.dpart[b * w.strideDp + k];
    if (k < jj) sW[k] = s; else sV[k - jj] = s;
  }
  __syncthreads();

  const int64_t rr = blockIdx.x * (int64_t)64 + lane;
  const int64_t r = j + 1 + rr;
  cplx<T> acc{0, 0};
  if (rr < m) {
    for (int k = sub; k < jj; k += 4) {
      acc = cadd(acc, cmul(V[r + (p + k) * n], sW[k]));
      acc = cadd(acc, cmul(weff(w, b, r, k, p, scal), sV[k]));
    }
  }
  red[sub][lane] = acc;
  __syncthreads();
  double pr = 0, pi = 0;
  if (sub == 0 && rr < m) {
    cplx<T> y = w.ypart[b * w.strideY + r];
    y = csub(y, cadd(cadd(red[0][lane], red[1][lane]), cadd(red[2][lane], red[3][lane])));
    y = cmul(tau, y);
    w.W[b * w.strideW + r + (int64_t)jj * n] = y;
    const cplx<T> v = V[r + j * n];
    const cplx<T> t = cmulc(y, v);
    pr = t.x; pi = t.y;
  }
  // the last workgroup only needs the atomically accumulated w^H v
  pr = block_sum(pr, sh);
  pi = block_sum(pi, sh);
  if (threadIdx.x == 0) {
    atomic_add_relaxed(&scal[S_YV], pr);
    atomic_add_relaxed(&scal[S_YV + 1], pi);
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
      hipLaunchKernelGGL(k_col<T>, dim3((unsigned)ceil_div(n - j, 64), 1, B), dim3(256), 0, s,
                         w, j, jj, p);
      PPH_CHECK_LAUNCH();
      const int64_t nt = ceil_div(m, 64);
      const int seg = kHemvSeg;
      hipLaunchKernelGGL(k_hemv_sym<T>, dim3((unsigned)nt, (unsigned)ceil_div(nt, seg), B), dim3(256),
                         0, s, w, j, jj, p, seg);
      PPH_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_fin<T>, dim3((unsigned)ceil_div(m, 64), 1, B), dim3(256), 0, s,
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
    g.lower = 1;              // only the lower triangle of the trailing matrix is read later
    launch_gemm<T>(g, s);
  }
  hipLaunchKernelGGL(k_last_diag<T>, dim3((unsigned)ceil_div(w.batch, 64)), dim3(64), 0, s, w);
  PPH_CHECK_LAUNCH();
}

template void tridiagonalise<double>(const TridiagWork<double>&, hipStream_t);
template void tridiagonalise<float>(const TridiagWork<float>&, hipStream_t);

}  // namespace pph
