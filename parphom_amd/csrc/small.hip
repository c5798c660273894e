// Hermitian -> real symmetric tridiagonal reduction of small matrices: one
// workgroup per matrix holds its whole matrix in LDS and runs every column of
// the reduction in one launch (no per-column kernel chain), so a batch of many
// small dynamical matrices (a few-atom cell on a dense q grid) costs one
// launch instead of three per column.
//
// Unblocked Householder reduction on the full Hermitian matrix (the ZHETD2
// steps, lower variant): for column j,
//   reflector H = I - tau v v^H from A(j+1:, j) (make_reflector: ZLARFG
//   formulas), v = (1, x / (alpha - beta)),
//   p = tau S v with S = A(j+1:, j+1:),
//   w = p - (tau/2) (p^H v) v,
//   S -= v w^H + w v^H.
// The outputs are the ones the blocked one-stage reduction writes, in the same
// layout and conventions, so the tridiagonal eigensolver and the
// back-transformation run unchanged after it: d (diagonal), e (real
// off-diagonal beta_j), tau_j, and V with column j = v_j explicitly (zero
// above row j + 1, one at row j + 1) plus zero columns up to the V stride.
// Every sum runs in a fixed order (row partials over fixed column classes,
// wave sums in a fixed butterfly), so results repeat bit for bit.
//
// This is the reduction half of pzheevx (diagonalize.F90:73-96) for orders up
// to kSmallMaxN (the matrix in LDS: 96 for complex128, 128 for complex64).
#include "common.hpp"
#include "kernels.hpp"

namespace pph {

constexpr int kSmallThreads = 512;

// NM: the LDS matrix's leading dimension (n <= NM)
template <typename T, int NM>
__global__ void __launch_bounds__(kSmallThreads) k_small_hetrd(TridiagWork<T> w, int64_t vcols) {
  constexpr int NT = kSmallThreads;
  constexpr int NP = 4;                       // column classes of the mat-vec partials
  __shared__ cplx<T> S[NM * NM];              // column-major, S(r, c) = S[r + c NM]
  __shared__ cplx<T> vv[NM], ww[NM], pp[NP][NM];
  const int64_t b = blockIdx.x;
  const int n = (int)w.n;
  const int tid = threadIdx.x;
  const cplx<T>* A = w.A + b * w.strideA;
  cplx<T>* V = w.V + b * w.strideV;
  T* d = w.d + b * w.strideDE;
  T* e = w.e + b * w.strideDE;
  cplx<T>* tau = w.tau + b * w.strideTau;

  // the lower triangle of A (coalesced reads), mirrored into the upper one
  for (int k = tid; k < n * n; k += NT) {
    const int r = k % n, c = k / n;
    if (r < c) continue;
    cplx<T> a = A[r + (int64_t)c * w.lda];
    if (r == c) a.y = 0;
    S[r + c * NM] = a;
    S[c + r * NM] = cconj(a);
  }
  // V columns n - 1 .. vcols - 1 carry no reflector: zero
  for (int64_t k = tid; k < (vcols - (n - 1)) * n; k += NT) V[(int64_t)(n - 1) * n + k] = cplx<T>{0, 0};
  __syncthreads();

  // Per column four barriers: wave 0 forms the reflector (its norm and the
  // dot p^H v are wave sums, m < 128 entries), all waves form the mat-vec
  // partials and the rank-2 update.
  Reflector rf{};
  for (int j = 0; j < n - 1; ++j) {
    const int o = j + 1, m = n - o;           // the trailing matrix S(o:, o:) has order m
    if (tid < WAVE) {
      double x2 = 0.0;
      for (int r = o + 1 + tid; r < n; r += WAVE) x2 += (double)cabs2(S[r + j * NM]);
      x2 = wave_sum(x2);
      const cplx<T> al = S[o + j * NM];
      rf = make_reflector(x2, (double)al.x, (double)al.y);
      if (tid == 0) {
        d[j] = S[j + j * NM].x;
        e[j] = (T)rf.beta;
        tau[j] = cplx<T>{(T)rf.tr, (T)rf.ti};
      }
      // v = (1, scale x) into LDS and into V(:, j), zero above row j + 1
      const cplx<T> sc{(T)rf.sr, (T)rf.si};
      for (int r = tid; r < n; r += WAVE) {
        cplx<T> v{0, 0};
        if (r == o) v = cplx<T>{1, 0};
        else if (r > o) v = cmul(sc, S[r + j * NM]);
        if (r >= o) vv[r - o] = v;
        V[r + (int64_t)j * n] = v;
      }
    }
    __syncthreads();
    // S v: NP partial sums per row, partial h over the columns c = h mod NP
    for (int u = tid; u < NP * m; u += NT) {
      const int r = u % m, h = u / m;
      const cplx<T>* col = S + (o + r) + o * NM;
      cplx<T> acc{0, 0};
      for (int c = h; c < m; c += NP) acc = cadd(acc, cmul(col[c * NM], vv[c]));
      pp[h][r] = acc;
    }
    __syncthreads();
    if (tid < WAVE) {
      // p = tau (sum of the partials), p^H v, then w = p + alpha2 v with
      // alpha2 = -tau/2 (p^H v)
      const cplx<T> tv{(T)rf.tr, (T)rf.ti};
      double yr = 0.0, yi = 0.0;
      for (int r = tid; r < m; r += WAVE) {
        cplx<T> y = pp[0][r];
#pragma unroll
        for (int h = 1; h < NP; ++h) y = cadd(y, pp[h][r]);
        y = cmul(tv, y);
        ww[r] = y;
        const cplx<T> v = vv[r];
        yr += (double)y.x * v.x + (double)y.y * v.y;
        yi += (double)y.x * v.y - (double)y.y * v.x;
      }
      yr = wave_sum(yr);
      yi = wave_sum(yi);
      const cplx<T> a2{(T)(-0.5 * (rf.tr * yr - rf.ti * yi)), (T)(-0.5 * (rf.tr * yi + rf.ti * yr))};
      for (int r = tid; r < m; r += WAVE) ww[r] = cadd(ww[r], cmul(a2, vv[r]));
    }
    __syncthreads();
    // S -= v w^H + w v^H (the diagonal kept real)
    for (int k = tid; k < m * m; k += NT) {
      const int r = k % m, c = k / m;
      cplx<T>& s = S[(o + r) + (o + c) * NM];
      cplx<T> t = csub(s, cadd(cmul(vv[r], cconj(ww[c])), cmul(ww[r], cconj(vv[c]))));
      if (r == c) t.y = 0;
      s = t;
    }
    __syncthreads();
  }
  if (tid == 0) {
    d[n - 1] = S[(n - 1) + (n - 1) * NM].x;
    e[n - 1] = 0;
    tau[n - 1] = cplx<T>{0, 0};
  }
}

template <typename T>
void tridiagonalise_small(const TridiagWork<T>& w, int64_t vcols, hipStream_t s) {
  const int64_t n = w.n;
  if (n < 1 || n > kSmallMaxN<T>)
    throw std::runtime_error("tridiagonalise_small: order " + std::to_string(n) + " outside 1.." +
                             std::to_string(kSmallMaxN<T>));
  if (vcols < n || w.strideV < n * vcols)
    throw std::runtime_error("tridiagonalise_small: V holds fewer than n columns per matrix");
  const dim3 grid((unsigned)w.batch), block(kSmallThreads);
  // the LDS matrix sized to the order (smaller tiles leave room for more
  // workgroups per CU)
  if (n <= 32) hipLaunchKernelGGL((k_small_hetrd<T, 32>), grid, block, 0, s, w, vcols);
  else if (n <= 64) hipLaunchKernelGGL((k_small_hetrd<T, 64>), grid, block, 0, s, w, vcols);
  else if (n <= 96) hipLaunchKernelGGL((k_small_hetrd<T, 96>), grid, block, 0, s, w, vcols);
  else if constexpr (kSmallMaxN<T> > 96)
    hipLaunchKernelGGL((k_small_hetrd<T, kSmallMaxN<T>>), grid, block, 0, s, w, vcols);
  PPH_CHECK_LAUNCH();
}

template void tridiagonalise_small<double>(const TridiagWork<double>&, int64_t, hipStream_t);
template void tridiagonalise_small<float>(const TridiagWork<float>&, int64_t, hipStream_t);

}  // namespace pph
