// Two-stage Hermitian -> real tridiagonal reduction, batched (the reduction
// half of what pzheevx does, diagonalize.F90:73-96, done the way a GPU with
// matrix cores wants it):
//
//   he2hb  dense -> band of half-bandwidth bw.  Per panel of bw columns:
//          Householder QR of the panel below the band (k_panel_qr, one
//          workgroup per matrix), its compact-WY factor T, then the two-sided
//          update of the trailing matrix
//              VT = V T,  X = A VT,  Mx = VT^H X,  Y = X - V Mx / 2,
//              A <- A - Y V^H - V Y^H
//          as GEMMs on the matrix cores (three real products per complex one
//          in FP32, four in FP64), on the lower triangle only (X from two
//          triangular products).  The trailing matrix is read about three
//          times per panel of bw columns instead of once per column.
//   hb2st  band -> real tridiagonal by bulge chasing (k_hb2st): sweep j
//          annihilates column j below the subdiagonal with a reflector of
//          length <= bw and chases the bulge it creates down the band, one
//          bw x bw block per step.  One workgroup per matrix; its waves run
//          consecutive sweeps, each at least three steps behind the one
//          before it (the regions of step k of sweep j and of steps > k + 2 of
//          sweep j - 1 are disjoint), with progress counters in LDS.
//   Q2     (apply_q2) the bulge-chasing reflectors applied to the tridiagonal
//          eigenvectors in blocks of bw sweeps x one step (compact WY), one
//          workgroup per column strip of Z walking the blocks in order.
//
// Q1 (the he2hb reflectors) is applied by back_transform with shift = bw.
#include "common.hpp"
#include "kernels.hpp"
#include "../../include/parphom_hip.h"

namespace pph {
// he2hb GEMMs with four real products per complex one (FP64) or three (FP32:
// the three-product form's cancellation in the imaginary part loosens the
// FP32 reduction by a small factor, which the FP64 refinement of the
// eigenpairs that follows removes)
template <typename T>
constexpr int kHe2hbFourM = sizeof(T) == 8 ? 1 : 0;
}  // namespace pph

namespace pph {

// ---------------------------------------------------------------- panel QR
// Householder QR of the panel P = A(r0:n, p0:p0+bw), r0 = p0 + bw, in place:
// R in its upper triangle, zeros below; the reflectors (explicit, unit
// first) into V(r0:, p0 + i), tau into tau[p0 + i].  One workgroup of 16
// waves per matrix; column c belongs to wave c % 16, which forms its dot
// with the reflector and updates it, and the owner of the next column also
// forms that column's alpha and |x|^2 for the next reflector.
constexpr int kQrThreads = 1024;

template <typename T>
__global__ void __launch_bounds__(kQrThreads) k_panel_qr(TwoStageWork<T> w, int64_t p0) {
  const int64_t b = blockIdx.x;
  const int64_t n = w.n, bw = w.bw, r0 = p0 + bw, m = n - r0;
  const int64_t lda = w.lda;
  cplx<T>* P = w.A + b * w.strideA + r0 + p0 * lda;
  cplx<T>* V = w.V + b * w.strideV + r0 + p0 * n;
  cplx<T>* tau = w.tau + b * w.strideTau + p0;
  const int tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
  constexpr int NWV = kQrThreads / WAVE;
  __shared__ double sh[NWV];
  __shared__ double s_next[3];           // alpha (re, im) and |x|^2 of the next column
  const int kq = (int)min<int64_t>(bw, m);
  {
    double s = 0.0;
    for (int64_t r = 1 + tid; r < m; r += kQrThreads) s += (double)cabs2(P[r]);
    s = block_sum(s, sh);
    if (tid == 0) {
      s_next[0] = (double)P[0].x;
      s_next[1] = (double)P[0].y;
      s_next[2] = s;
    }
  }
  for (int i = 0; i < kq; ++i) {
    __syncthreads();
    const Reflector rf = make_reflector(s_next[2], s_next[0], s_next[1]);
    const cplx<T> sc{(T)rf.sr, (T)rf.si};
    cplx<T>* Pi = P + (int64_t)i * lda;
    cplx<T>* Vi = V + (int64_t)i * n;
    for (int64_t r = i + tid; r < m; r += kQrThreads) {
      if (r == i) {
        Vi[r] = cplx<T>{1, 0};
        Pi[r] = cplx<T>{(T)rf.beta, 0};
      } else {
        Vi[r] = cmul(Pi[r], sc);
        Pi[r] = cplx<T>{0, 0};
      }
    }
    if (tid == 0) tau[i] = cplx<T>{(T)rf.tr, (T)rf.ti};
    __syncthreads();                       // v visible to every wave
    // conj(tau) for H^H
    const double ctr = rf.tr, cti = -rf.ti;
    for (int c = wave; c < bw; c += NWV) {
      if (c <= i) continue;
      cplx<T>* Pc = P + (int64_t)c * lda;
      // QU rows per lane in flight per iteration (from clamped addresses, so
      // the loads are unconditional and issued together)
      constexpr int QU = 8;
      double zr = 0.0, zi = 0.0;
      for (int64_t r0_ = i; r0_ < m; r0_ += QU * WAVE) {
        cplx<T> vv[QU], xx[QU];
#pragma unroll
        for (int u = 0; u < QU; ++u) {
          const int64_t r = min(r0_ + lane + u * WAVE, m - 1);
          vv[u] = Vi[r];
          xx[u] = Pc[r];
        }
#pragma unroll
        for (int u = 0; u < QU; ++u) {
          // conj(v) x
          const double ok = (r0_ + lane + u * WAVE < m) ? 1.0 : 0.0;
          zr += ok * ((double)vv[u].x * xx[u].x + (double)vv[u].y * xx[u].y);
          zi += ok * ((double)vv[u].x * xx[u].y - (double)vv[u].y * xx[u].x);
        }
      }
      zr = wave_sum(zr);
      zi = wave_sum(zi);
      // P(:, c) -= conj(tau) z v
      const cplx<T> f{(T)(ctr * zr - cti * zi), (T)(ctr * zi + cti * zr)};
      const bool next = (c == i + 1);
      double s = 0.0;
      for (int64_t r0_ = i; r0_ < m; r0_ += QU * WAVE) {
        cplx<T> vv[QU], xx[QU];
#pragma unroll
        for (int u = 0; u < QU; ++u) {
          const int64_t r = min(r0_ + lane + u * WAVE, m - 1);
          vv[u] = Vi[r];
          xx[u] = Pc[r];
        }
#pragma unroll
        for (int u = 0; u < QU; ++u) {
          const int64_t r = r0_ + lane + u * WAVE;
          if (r >= m) continue;
          const cplx<T> x = csub(xx[u], cmul(f, vv[u]));
          Pc[r] = x;
          if (next) {
            if (r > i + 1) s += (double)cabs2(x);
            else if (r == i + 1) { s_next[0] = (double)x.x; s_next[1] = (double)x.y; }
          }
        }
      }
      if (next) {
        s = wave_sum(s);
        if (lane == 0) s_next[2] = s;
      }
    }
  }
}

// ---------------------------------------------------------------- tall Gram products
// Partial products C = A(rows)^H B(rows) over chunks of the m rows: A and B
// m x bw (leading dimension ld), one workgroup per (chunk, matrix); partials
// at gpart[(b * nchunk + chunk) * bw^2 + i + j bw].  The chunks are summed in a
// fixed order by the consumer, so the result does not depend on scheduling.
template <typename T>
__global__ void __launch_bounds__(256) k_tall_gram(const cplx<T>* A, const cplx<T>* Bm, int64_t ld,
                                                   int64_t strideAB, int64_t m, int bw,
                                                   int64_t nchunk, cplx<T>* gpart) {
  __shared__ cplx<T> As[32][33], Bs[32][33];
  const int64_t chunk = blockIdx.x, b = blockIdx.y;
  const int64_t rows = ceil_div_dev(m, nchunk);
  const int64_t rb = chunk * rows, re = min(m, rb + rows);
  const cplx<T>* Ab = A + b * strideAB;
  const cplx<T>* Bb = Bm + b * strideAB;
  const int tid = threadIdx.x;
  const int oi = tid % 32, oj0 = tid / 32;          // outputs (oi, oj0 + 8 q), q < 4
  double accr[4] = {0, 0, 0, 0}, acci[4] = {0, 0, 0, 0};
  for (int64_t t0 = rb; t0 < re; t0 += 32) {
    __syncthreads();
    for (int e = tid; e < 32 * 32; e += 256) {
      const int r = e % 32, c = e / 32;
      const int64_t gr = t0 + r;
      const bool ok = gr < re && c < bw;
      As[r][c] = ok ? Ab[gr + c * ld] : cplx<T>{0, 0};
      Bs[r][c] = ok ? Bb[gr + c * ld] : cplx<T>{0, 0};
    }
    __syncthreads();
#pragma unroll 4
    for (int r = 0; r < 32; ++r) {
      const cplx<T> a = As[r][oi];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const cplx<T> bb = Bs[r][oj0 + 8 * q];
        accr[q] += (double)a.x * bb.x + (double)a.y * bb.y;
        acci[q] += (double)a.x * bb.y - (double)a.y * bb.x;
      }
    }
  }
  cplx<T>* out = gpart + (b * nchunk + chunk) * (int64_t)bw * bw;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int oj = oj0 + 8 * q;
    if (oi < bw && oj < bw) out[oi + oj * bw] = cplx<T>{(T)accr[q], (T)acci[q]};
  }
}

template <typename T>
static void tall_gram(const cplx<T>* A, const cplx<T>* Bm, int64_t ld, int64_t strideAB, int64_t m,
                      int bw, int64_t nchunk, int64_t batch, cplx<T>* gpart, hipStream_t s) {
  hipLaunchKernelGGL(k_tall_gram<T>, dim3((unsigned)nchunk, (unsigned)batch), dim3(256), 0, s, A, Bm,
                     ld, strideAB, m, bw, nchunk, gpart);
  PPH_CHECK_LAUNCH();
}

// C(bw x bw, leading dimension bw) = sum of the chunks' partials, in order
template <typename T>
__global__ void __launch_bounds__(1024) k_gram_sum(const cplx<T>* gpart, int64_t nchunk, int bw,
                                                   cplx<T>* C, int64_t strideC) {
  const int64_t b = blockIdx.x;
  const int e = threadIdx.x;
  if (e >= bw * bw) return;
  double sr = 0.0, si = 0.0;
  const cplx<T>* g = gpart + b * nchunk * (int64_t)bw * bw + e;
  for (int64_t c = 0; c < nchunk; ++c) {
    const cplx<T> v = g[c * (int64_t)bw * bw];
    sr += (double)v.x;
    si += (double)v.y;
  }
  C[b * strideC + e] = cplx<T>{(T)sr, (T)si};
}

// T of the panel (forward compact WY, H_0 ... H_{bw-1} = I - V T V^H) from
// G = V^H V (the chunk partials) and tau: T(i, i) = tau_i,
// T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i).  One workgroup per matrix.
template <typename T>
__global__ void __launch_bounds__(64) k_panel_t(const cplx<T>* gpart, int64_t nchunk, int bw,
                                                const cplx<T>* tau, int64_t strideTau, int64_t p0,
                                                cplx<T>* Tp, int64_t strideTp) {
  __shared__ cplx<double> G[32][33], Tm[32][33];
  const int64_t b = blockIdx.x;
  const int tid = threadIdx.x;
  for (int e = tid; e < bw * bw; e += 64) {
    double sr = 0.0, si = 0.0;
    const cplx<T>* g = gpart + b * nchunk * (int64_t)bw * bw + e;
    for (int64_t c = 0; c < nchunk; ++c) {
      const cplx<T> v = g[c * (int64_t)bw * bw];
      sr += (double)v.x;
      si += (double)v.y;
    }
    G[e % bw][e / bw] = cplx<double>{sr, si};
    Tm[e % bw][e / bw] = cplx<double>{0, 0};
  }
  __syncthreads();
  const cplx<T>* tb = tau + b * strideTau + p0;
  for (int i = 0; i < bw; ++i) {
    const cplx<double> ti{(double)tb[i].x, (double)tb[i].y};
    cplx<double> v{0, 0};
    if (tid < i) {
      for (int l = tid; l < i; ++l) v = cadd(v, cmul(Tm[tid][l], G[l][i]));
      v = cmul(cplx<double>{-ti.x, -ti.y}, v);
    }
    __syncthreads();
    if (tid < i) Tm[tid][i] = v;
    if (tid == i) Tm[i][i] = ti;
    __syncthreads();
  }
  cplx<T>* out = Tp + b * strideTp;
  for (int e = tid; e < bw * bw; e += 64)
    out[e] = cplx<T>{(T)Tm[e % bw][e / bw].x, (T)Tm[e % bw][e / bw].y};
}

// ---------------------------------------------------------------- band extraction
// AB[c * 2bw + d] = A(c + d, c) for d <= bw, zero for bw < d < 2bw
template <typename T>
__global__ void __launch_bounds__(256) k_extract_band(const cplx<T>* A, int64_t lda, int64_t strideA,
                                                      int64_t n, int bw, cplx<T>* AB,
                                                      int64_t strideAB) {
  const int ldb = 2 * bw;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t b = blockIdx.y;
  if (e >= n * ldb) return;
  const int64_t c = e / ldb;
  const int d = (int)(e % ldb);
  cplx<T> v{0, 0};
  if (d <= bw && c + d < n) v = A[b * strideA + (c + d) + c * lda];
  if (d == 0) v.y = T(0);
  AB[b * strideAB + e] = v;
}

template <typename T>
void he2hb(const TwoStageWork<T>& w, hipStream_t s) {
  const int64_t n = w.n, bw = w.bw, B = w.batch;
  if (bw != 16 && bw != 32) throw std::runtime_error("he2hb: band width must be 16 or 32");
  launch_fill_zero(w.V, B * w.strideV * (int64_t)sizeof(cplx<T>), s);
  launch_fill_zero(w.tau, B * w.strideTau * (int64_t)sizeof(cplx<T>), s);
  for (int64_t p0 = 0; p0 + bw < n; p0 += bw) {
    const int64_t r0 = p0 + bw, m = n - r0;
    hipLaunchKernelGGL(k_panel_qr<T>, dim3((unsigned)B), dim3(kQrThreads), 0, s, w, p0);
    PPH_CHECK_LAUNCH();
    const int64_t nch = std::max<int64_t>(1, std::min<int64_t>(w.nchunk, ceil_div(m, 256)));
    const cplx<T>* Vp = w.V + r0 + p0 * n;
    // T of the panel from G = V^H V
    tall_gram<T>(Vp, Vp, n, w.strideV, m, (int)bw, nch, B, w.gpart, s);
    hipLaunchKernelGGL(k_panel_t<T>, dim3((unsigned)B), dim3(64), 0, s, w.gpart, nch, (int)bw, w.tau,
                       w.strideTau, p0, w.Tp, w.strideTp);
    PPH_CHECK_LAUNCH();
    // VT = V T   (m x bw)
    {
      GemmParams<T> g{};
      g.opA = 'N'; g.opB = 'N';
      g.m = m; g.n = bw; g.k = bw;
      g.alpha = cplx<T>{1, 0}; g.beta = cplx<T>{0, 0};
      g.A = Vp; g.lda = n; g.strideA = w.strideV;
      g.B = w.Tp; g.ldb = bw; g.strideB = w.strideTp;
      g.C = w.VT + r0; g.ldc = n; g.strideC = w.strideXV;
      g.batch = B;
      g.four_m = kHe2hbFourM<T>;
      launch_gemm<T>(g, s);
    }
    // X = A22 VT   (m x bw), from the lower triangle of A22: its lower part
    // with the diagonal, then the strict lower part conjugate-transposed
    {
      GemmParams<T> g{};
      g.opA = 'N'; g.opB = 'N';
      g.m = m; g.n = bw; g.k = m;
      g.alpha = cplx<T>{1, 0}; g.beta = cplx<T>{0, 0};
      g.A = w.A + r0 + r0 * w.lda; g.lda = w.lda; g.strideA = w.strideA;
      g.B = w.VT + r0; g.ldb = n; g.strideB = w.strideXV;
      g.C = w.X + r0; g.ldc = n; g.strideC = w.strideXV;
      g.batch = B;
      g.four_m = kHe2hbFourM<T>;
      g.tri = 1;
      launch_gemm<T>(g, s);
      g.opA = 'C'; g.tri = 2; g.beta = cplx<T>{1, 0};
      launch_gemm<T>(g, s);
    }
    // Mx = VT^H X = T^H V^H A22 V T   (bw x bw, into Tp + bw^2)
    cplx<T>* Mx = w.Tp + bw * bw;
    tall_gram<T>(w.VT + r0, w.X + r0, n, w.strideXV, m, (int)bw, nch, B, w.gpart, s);
    hipLaunchKernelGGL(k_gram_sum<T>, dim3((unsigned)B), dim3(1024), 0, s, w.gpart, nch, (int)bw, Mx,
                       w.strideTp);
    PPH_CHECK_LAUNCH();
    // Y = X - V Mx / 2   (in X)
    {
      GemmParams<T> g{};
      g.opA = 'N'; g.opB = 'N';
      g.m = m; g.n = bw; g.k = bw;
      g.alpha = cplx<T>{T(-0.5), 0}; g.beta = cplx<T>{1, 0};
      g.A = Vp; g.lda = n; g.strideA = w.strideV;
      g.B = Mx; g.ldb = bw; g.strideB = w.strideTp;
      g.C = w.X + r0; g.ldc = n; g.strideC = w.strideXV;
      g.batch = B;
      g.four_m = kHe2hbFourM<T>;
      launch_gemm<T>(g, s);
    }
    // A22 -= Y V^H + V Y^H   (the lower triangle: tiles wholly above the
    // diagonal are skipped, and nothing above the diagonal is read later)
    {
      GemmParams<T> g{};
      g.opA = 'N'; g.opB = 'C';
      g.m = m; g.n = m; g.k = 2 * bw; g.k1 = bw;
      g.alpha = cplx<T>{-1, 0}; g.beta = cplx<T>{1, 0};
      g.A = w.X + r0; g.lda = n; g.strideA = w.strideXV;
      g.B = Vp; g.ldb = n; g.strideB = w.strideV;
      g.A2 = Vp; g.strideA2 = w.strideV;
      g.B2 = w.X + r0; g.strideB2 = w.strideXV;
      g.C = w.A + r0 + r0 * w.lda; g.ldc = w.lda; g.strideC = w.strideA;
      g.batch = B;
      g.four_m = kHe2hbFourM<T>;
      g.lower = 1;
      launch_gemm<T>(g, s);
    }
  }
  const int64_t ldb = 2 * bw;
  hipLaunchKernelGGL(k_extract_band<T>, dim3((unsigned)ceil_div(n * ldb, 256), (unsigned)B),
                     dim3(256), 0, s, w.A, w.lda, w.strideA, n, (int)bw, w.AB, w.strideAB);
  PPH_CHECK_LAUNCH();
}

template void he2hb<float>(const TwoStageWork<float>&, hipStream_t);
template void he2hb<double>(const TwoStageWork<double>&, hipStream_t);

// ---------------------------------------------------------------- bulge chasing
// Step k of sweep j works on the row range R_k = [s, s + L), s = j + 1 + k bw:
//   the reflector H from the first column of X = A(R_k, R_{k-1}) (k = 0: from
//   column j), X <- H^H X (its first column becomes (beta, 0, ...)),
//   D = A(R_k, R_k) <- H^H D H, and the block below, Y = A(R_{k+1}, R_k) <- Y H,
//   which is the next step's X.
// Band element (r, c), r >= c, sits at AB[c * 2bw + r - c]: band column c is
// 2bw contiguous entries.  A step loads its band columns R_k whole (one
// coalesced load per column, lane = entry) into registers, finishes the
// previous step's columns (its X = the previous Y lives there) in the wave's
// LDS buffer and stores them whole, then moves the new columns into the
// buffer and updates D and Y there.  So every band column is loaded and
// stored once per sweep, with whole-line accesses; the published progress
// counts the steps whose columns are stored.
template <typename T, int BW>
__device__ __forceinline__ T sum_halves(T v) {      // over the lanes of one column / row
#pragma unroll
  for (int o = BW; o < WAVE; o <<= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

template <typename T, int BW, int NW>
__global__ void __launch_bounds__(WAVE * NW) k_hb2st(TwoStageWork<T> w) {
  constexpr int NG = WAVE / BW;
  constexpr int RPL = BW / NG;             // rows (column layout) or columns (row layout) per lane
  constexpr int LDB = 2 * BW;
  constexpr int CPI = WAVE / LDB;          // band columns per load instruction (1 or 2)
  constexpr int NLD = BW / CPI;            // load instructions per step
  constexpr int SH = 4096;                 // progress = sweep * SH + steps stored
  const int64_t b = blockIdx.x;
  const int64_t n = w.n;
  cplx<T>* AB = w.AB + b * w.strideAB;
  cplx<T>* V2 = w.V2 + b * w.strideV2;
  cplx<T>* tau2 = w.tau2 + b * w.strideTau2;
  const int64_t kmax = w.kmax;
  const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
  const int cl = lane % BW, hl = lane / BW, q0 = hl * RPL;
  const int dl = lane % LDB, il = lane / LDB;          // band entry, column offset of a load
  __shared__ int prog[NW];
  // band columns of the step, [column][entry]; a pitch of 2bw + 2 makes both the
  // walk along a column (lane = entry) and the walk along a diagonal (lane =
  // column, entry - column fixed) free of LDS bank conflicts
  __shared__ cplx<T> buf[NW][BW][LDB + 2];
  __shared__ cplx<T> sv[NW][BW], sw[NW][BW];
  __shared__ int stalled;
  if (threadIdx.x < NW) prog[threadIdx.x] = -1;
  if (threadIdx.x == 0) stalled = 0;
  __syncthreads();
  cplx<T>(*B)[LDB + 2] = buf[wave];
  auto wave_sync = []() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  auto publish = [&](int64_t j, int64_t done) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane == 0)
      __hip_atomic_store(&prog[wave], (int)(j * SH + done), __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
  };
  // store the buffer's first ncol columns as band columns c0.. (whole columns)
  auto store_cols = [&](int c0, int ncol) {
#pragma unroll
    for (int q = 0; q < NLD; ++q) {
      const int i = q * CPI + il;
      if (i < ncol) AB[(c0 + i) * LDB + dl] = B[i][dl];
    }
  };
  const int pw = (wave + NW - 1) % NW;
  for (int64_t j = wave; j < n - 1; j += NW) {
    const int64_t K = hb2st_tasks(n, BW, j);
    const int64_t Kp = (j > 0) ? hb2st_tasks(n, BW, j - 1) : 0;
    const int ji = (int)j;
    for (int64_t k = 0; k < K; ++k) {
      if (j > 0) {
        // sweep j - 1 must have stored the columns of its steps 0 .. k + 2.
        // The wait is bounded (about 0.4 s, far beyond any step): a stall is
        // flagged in info and every later wait of the workgroup is skipped,
        // so the kernel always ends
        const int need = (int)((j - 1) * SH + min<int64_t>(k + 3, Kp));
        int spins = 0;
        while (__hip_atomic_load(&prog[pw], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < need &&
               __hip_atomic_load(&stalled, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0) {
          __builtin_amdgcn_s_sleep(1);
          if (++spins > (1 << 23)) {
            if (lane == 0)
              __hip_atomic_store(&stalled, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            break;
          }
        }
      }
      const int si = (int)(j + 1 + k * BW);             // 32-bit offsets into the band
      const int L = (int)min<int64_t>(BW, n - si);
      const bool below = k + 1 < K;
      const int L2 = below ? (int)min<int64_t>(BW, n - si - BW) : 0;
      // this step's band columns si .. si + L - 1, whole
      // FP32: loaded into registers now, so the loads overlap the left update;
      // FP64 (twice the registers): loaded after the previous columns are stored
      constexpr bool early = sizeof(T) == 4;
      cplx<T> R[NLD];
      if (early) {
#pragma unroll
        for (int q = 0; q < NLD; ++q) {
          const int i = q * CPI + il;
          R[q] = AB[(si + min(i, L - 1)) * LDB + dl];  // columns past the matrix: never used
        }
      }
      if (k == 0) {
        // X = column j: band entries 1 .. L of column j, into buffer column 0
        if (lane < LDB) B[0][lane] = AB[ji * LDB + lane];
        wave_sync();
      }
      // X(r, i) = B[i][xo + r - i]: the previous step's Y (xo = bw), or column j (xo = 1)
      const int xo = (k == 0) ? 1 : BW;
      const int xcols = (k == 0) ? 1 : BW;
      // reflector from column 0 of X
      double nn = 0.0;
      {
        const int r = lane;
        if (r >= 1 && r < L) {
          const cplx<T> x = B[0][xo + r];
          nn = (double)x.x * x.x + (double)x.y * x.y;
        }
      }
      nn = wave_sum(nn);
      const cplx<T> x0 = B[0][xo];
      const Reflector rf = make_reflector(nn, (double)x0.x, (double)x0.y);
      const cplx<T> tau{(T)rf.tr, (T)rf.ti};
      if (lane < BW) {
        const int r = lane;
        cplx<T> v{0, 0};
        if (r == 0) v = cplx<T>{1, 0};
        else if (r < L) {
          const cplx<T> x = B[0][xo + r];
          v = cplx<T>{(T)(x.x * rf.sr - x.y * rf.si), (T)(x.x * rf.si + x.y * rf.sr)};
        }
        sv[wave][r] = v;
      }
      wave_sync();
      cplx<T> vr[RPL];                                   // v at the lane's rows
#pragma unroll
      for (int i = 0; i < RPL; ++i) vr[i] = sv[wave][q0 + i];
      const cplx<T> vcl = sv[wave][cl];                   // v at the lane's column
      // X <- H^H X on columns 1 ..: z_c = sum_r conj(v_r) X(r, c),
      // X(r, c) -= conj(tau) v_r z_c (column layout, in the buffer)
      if (k >= 1) {
        // v_r = 0 for r >= L, so rows past the matrix neither contribute nor change;
        // column 0 is overwritten below
        cplx<T> xs[RPL];
#pragma unroll
        for (int i = 0; i < RPL; ++i) xs[i] = B[cl][BW + q0 + i - cl];
        cplx<T> z{0, 0};
#pragma unroll
        for (int i = 0; i < RPL; ++i) z = cadd(z, cmulc(vr[i], xs[i]));
        z.x = sum_halves<T, BW>(z.x);
        z.y = sum_halves<T, BW>(z.y);
        const cplx<T> f = cmul(cplx<T>{tau.x, -tau.y}, z);
#pragma unroll
        for (int i = 0; i < RPL; ++i) B[cl][BW + q0 + i - cl] = csub(xs[i], cmul(f, vr[i]));
      }
      // column 0 of X becomes (beta, 0, ...)
      if (lane < BW && lane < L) B[0][xo + lane] = (lane == 0) ? cplx<T>{(T)rf.beta, 0} : cplx<T>{0, 0};
      wave_sync();
      // the previous step's columns (or column j) are final: store them whole
      if (k == 0) {
        if (lane < LDB) AB[ji * LDB + lane] = B[0][lane];
      } else {
        store_cols(si - BW, xcols);
      }
      wave_sync();
      // this step's columns into the buffer
      if (!early) {
#pragma unroll
        for (int q = 0; q < NLD; ++q) {
          const int i = q * CPI + il;
          R[q] = AB[(si + min(i, L - 1)) * LDB + dl];
        }
      }
#pragma unroll
      for (int q = 0; q < NLD; ++q) B[q * CPI + il][dl] = R[q];
      wave_sync();
      // D <- H^H D H (column layout): y_c = sum_r conj(D(r, c)) v_r, w = tau y -
      // |tau|^2 (v^H y) v / 2, D -= w v^H + v w^H (lower triangle kept)
      cplx<T> D[RPL];
#pragma unroll
      for (int i = 0; i < RPL; ++i) {
        const int r = q0 + i, c = cl;
        const bool lo = r >= c;
        const cplx<T> u = lo ? B[c][r - c] : B[r][c - r];
        const bool in = r < L && c < L;
        D[i] = cplx<T>{in ? u.x : T(0), in ? (lo ? u.y : -u.y) : T(0)};
      }
      cplx<T> y{0, 0};
#pragma unroll
      for (int i = 0; i < RPL; ++i) y = cadd(y, cmulc(D[i], vr[i]));
      y.x = sum_halves<T, BW>(y.x);
      y.y = sum_halves<T, BW>(y.y);
      double hh = (hl == 0 && cl < L) ? (double)vcl.x * y.x + (double)vcl.y * y.y : 0.0;
      hh = wave_sum(hh);                                  // v^H D v (real)
      const double t2 = rf.tr * rf.tr + rf.ti * rf.ti;
      const cplx<T> tyv = cmul(tau, y);
      const T hs = (T)(0.5 * t2 * hh);
      const cplx<T> wc{tyv.x - hs * vcl.x, tyv.y - hs * vcl.y};
      if (hl == 0) sw[wave][cl] = wc;
      // Y <- Y H (row layout: t_r = sum_c Y(r, c) v_c), Y(r, c) = B[c][bw + r - c]
      // (v_c = 0 for c >= L; rows yr >= L2 are past the matrix and never stored)
      cplx<T> t{0, 0};
      cplx<T> ys[RPL];
#pragma unroll
      for (int i = 0; i < RPL; ++i) ys[i] = B[q0 + i][BW + cl - q0 - i];
      if (below) {
#pragma unroll
        for (int i = 0; i < RPL; ++i) t = cadd(t, cmul(ys[i], vr[i]));
        t.x = sum_halves<T, BW>(t.x);
        t.y = sum_halves<T, BW>(t.y);
      }
      wave_sync();
      cplx<T> wrs[RPL];
#pragma unroll
      for (int i = 0; i < RPL; ++i) wrs[i] = sw[wave][q0 + i];
#pragma unroll
      for (int i = 0; i < RPL; ++i) {
        const int r = q0 + i, c = cl;
        cplx<T> d = csub(D[i], cmulc(vcl, wrs[i]));        // w_r conj(v_c)
        d = csub(d, cmulc(wc, vr[i]));                     // v_r conj(w_c)
        if (r == c) d.y = T(0);
        if (r < L && c < L && r >= c) B[c][r - c] = d;
      }
      if (below) {
        const cplx<T> tt = cmul(tau, t);
#pragma unroll
        for (int i = 0; i < RPL; ++i)
          B[q0 + i][BW + cl - q0 - i] = csub(ys[i], cmulc(vr[i], tt));   // tau t_r conj(v_c)
      }
      // the reflector for the back-transformation
      if (hl == 0) V2[(j * kmax + k) * BW + cl] = sv[wave][cl];
      if (lane == 0) tau2[j * kmax + k] = tau;
      wave_sync();
      if (!below) {
        // the sweep's last columns are final too
        store_cols(si, L);
        publish(j, K);
      } else {
        publish(j, k);                                     // steps 0 .. k - 1 stored
      }
      wave_sync();
    }
  }
  __syncthreads();
  // the real tridiagonal
  T* d = w.d + b * w.strideDE;
  T* e = w.e + b * w.strideDE;
  for (int64_t c = threadIdx.x; c < n; c += blockDim.x) {
    d[c] = AB[c * LDB].x;
    e[c] = (c + 1 < n) ? AB[c * LDB + 1].x : T(0);
  }
  if (threadIdx.x == 0 && stalled && w.info) atomicOr(w.info + b, PPH_INFO_NOCONV);
}

template <typename T>
void hb2st(const TwoStageWork<T>& w, hipStream_t s) {
  const int64_t n = w.n, B = w.batch;
  if (n < 2) return;
  if (hb2st_tasks(n, w.bw, 0) >= 4096) throw std::runtime_error("hb2st: n too large for the band width");
  if (n * 2 * w.bw >= (int64_t(1) << 31)) throw std::runtime_error("hb2st: band too large for 32-bit offsets");
  constexpr int NW = sizeof(T) == 4 ? 8 : 4;      // the register budget of a wave: 256 or 512 VGPRs
  if (w.bw == 32)
    hipLaunchKernelGGL((k_hb2st<T, 32, NW>), dim3((unsigned)B), dim3(WAVE * NW), 0, s, w);
  else if (w.bw == 16)
    hipLaunchKernelGGL((k_hb2st<T, 16, NW>), dim3((unsigned)B), dim3(WAVE * NW), 0, s, w);
  else
    throw std::runtime_error("hb2st: band width must be 16 or 32");
  PPH_CHECK_LAUNCH();
}

template void hb2st<float>(const TwoStageWork<float>&, hipStream_t);
template void hb2st<double>(const TwoStageWork<double>&, hipStream_t);

// ---------------------------------------------------------------- Q2
// Block (g, k): the reflectors of sweeps j0 = g bw .. j0 + bw - 1 at step k,
// S = H(j0, k) H(j0 + 1, k) ... = I - V T V^H with V (2bw x bw) holding
// reflector i at local rows i .. i + bw - 1 (global row j0 + 1 + k bw + local).
// Q2 = prod_g prod_{k descending} S(g, k) (groups in order), so Q2 Z applies
// the groups from the last to the first and, within a group, k = 0, 1, ...
// (the reflectors of one sweep act on disjoint rows, and S(g, k) must precede
// S(g, k - 1) in the product since H(j, k) is formed before H(j + 1, k - 1)).

// V of block (g, k) into Vs (row-major: Vs[r][i]), zero where no reflector
template <typename T, int BW>
__device__ __forceinline__ void gather_v(const cplx<T>* V2, const cplx<T>* tau2, int64_t kmax,
                                         int64_t n, int64_t j0, int64_t k, cplx<T> (*Vs)[BW + 1],
                                         cplx<T>* ts) {
  for (int e = threadIdx.x; e < 2 * BW * BW; e += blockDim.x) {
    const int rr = e % (2 * BW), i = e / (2 * BW);
    const int64_t j = j0 + i;
    cplx<T> v{0, 0};
    const bool has = (j < n - 1) && (k < hb2st_tasks(n, BW, j));
    if (has && rr >= i && rr < i + BW) v = V2[(j * kmax + k) * BW + (rr - i)];
    Vs[rr][i] = v;
  }
  for (int i = threadIdx.x; i < BW; i += blockDim.x) {
    const int64_t j = j0 + i;
    const bool has = (j < n - 1) && (k < hb2st_tasks(n, BW, j));
    ts[i] = has ? tau2[j * kmax + k] : cplx<T>{0, 0};
  }
}

// Per block: V row-major (Vb[rr * bw + i]) and VT = V T column-major
// (VTb[rr + i * 2bw]), at block index (g kmax + k) * 2bw^2 of each.
// One workgroup per (k, g, matrix).
template <typename T, int BW>
__global__ void __launch_bounds__(256) k_q2_prep(Q2Work<T> w) {
  __shared__ cplx<T> Vs[2 * BW][BW + 1];
  __shared__ cplx<T> ts[BW];
  // G = V^H V strictly below the diagonal of M (G(l, i) at M[i][l], l < i),
  // T on and above it (T(l, i) at M[l][i]): the recurrence for T reads only
  // G(l, i) with l < i, so one array holds both
  __shared__ cplx<double> M[BW][BW + 1];
  const int64_t k = blockIdx.x, g = blockIdx.y, b = blockIdx.z;
  const int64_t n = w.n, j0 = g * BW;
  if (j0 >= n - 1 || k >= hb2st_tasks(n, BW, j0)) return;
  const int tid = threadIdx.x;
  gather_v<T, BW>(w.V2 + b * w.strideV2, w.tau2 + b * w.strideTau2, w.kmax, n, j0, k, Vs, ts);
  __syncthreads();
  const int64_t blk = (g * w.kmax + k) * (int64_t)(2 * BW * BW);
  cplx<T>* vout = w.Vb + b * w.strideVTb + blk;
  for (int e = tid; e < 2 * BW * BW; e += 256) vout[e] = Vs[e / BW][e % BW];
  // G = V^H V (column i nonzero on rows i .. i + bw - 1), entries a < c
  for (int e = tid; e < BW * BW; e += 256) {
    const int a = e % BW, c = e / BW;
    if (a >= c) continue;
    const int lo = max(a, c), hi = min(a, c) + BW;
    double sr = 0.0, si = 0.0;
    for (int rr = lo; rr < hi; ++rr) {
      const cplx<T> x = Vs[rr][a], y = Vs[rr][c];
      sr += (double)x.x * y.x + (double)x.y * y.y;
      si += (double)x.x * y.y - (double)x.y * y.x;
    }
    M[c][a] = cplx<double>{sr, si};
  }
  __syncthreads();
  // T column by column (T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i)): only the
  // first bw lanes work, so the first wave runs the recurrence alone and
  // synchronises at wave level; column i is read only after it is written
  if (tid < WAVE) {
    for (int i = 0; i < BW; ++i) {
      const cplx<double> ti{(double)ts[i].x, (double)ts[i].y};
      if (tid < i) {
        cplx<double> v{0, 0};
        for (int l = tid; l < i; ++l) v = cadd(v, cmul(M[tid][l], M[i][l]));
        M[tid][i] = cmul(cplx<double>{-ti.x, -ti.y}, v);
      }
      if (tid == i) M[i][i] = ti;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  __syncthreads();
  // VT(rr, c) = sum_{l <= c} V(rr, l) T(l, c)
  cplx<T>* out = w.VTb + b * w.strideVTb + blk;
  for (int e = tid; e < 2 * BW * BW; e += 256) {
    const int rr = e % (2 * BW), c = e / (2 * BW);
    double sr = 0.0, si = 0.0;
    for (int l = max(0, rr - BW + 1); l <= min(c, rr); ++l) {
      const cplx<T> x = Vs[rr][l];
      const cplx<double> t = M[l][c];
      sr += (double)x.x * t.x - (double)x.y * t.y;
      si += (double)x.x * t.y + (double)x.y * t.x;
    }
    out[rr + c * (2 * BW)] = cplx<T>{(T)sr, (T)si};
  }
}

// One 16 x 16 complex tile of C = A B on the matrix cores (3M: three real
// MFMAs per K step of 4), A(m, k) and B(k, n) read from LDS through the
// accessors; K steps [k0, k1).  Lane l holds column l % 16 of the tile, rows
// mfma<T>::row(l, r), r < 4.
template <typename T, typename FA, typename FB>
__device__ __forceinline__ void tile3m(FA&& A, FB&& B, int k0, int k1, cplx<T> out[4]) {
  using acc_t = typename mfma<T>::acc;
  const int lane = threadIdx.x % WAVE;
  acc_t p1{0, 0, 0, 0}, p2{0, 0, 0, 0}, p3{0, 0, 0, 0};
  for (int k = k0; k < k1; k += 4) {
    const cplx<T> a = A(lane % 16, k + lane / 16);
    const cplx<T> b = B(k + lane / 16, lane % 16);
    p1 = mfma<T>::op(a.x, b.x, p1);
    p2 = mfma<T>::op(a.y, b.y, p2);
    p3 = mfma<T>::op(a.x + a.y, b.x + b.y, p3);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) out[r] = cplx<T>{p1[r] - p2[r], p3[r] - p1[r] - p2[r]};
}

// Buffer loads and stores of one complex element through a raw buffer
// resource: an offset past the resource reads as zero and drops the store,
// so lanes outside the matrix need no branch.
typedef unsigned int q2u2 __attribute__((ext_vector_type(2)));
typedef unsigned int q2u4 __attribute__((ext_vector_type(4)));
constexpr unsigned kQ2Off = 0xFFFFFFFFu;         // an offset no resource covers
// bits of the resource's last word for a raw (untyped) buffer
constexpr int kRawBufferWord3 = 0x00020000;
__device__ __forceinline__ cplx<float> q2_load(__amdgpu_buffer_rsrc_t r, unsigned off, float) {
  return __builtin_bit_cast(cplx<float>, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ cplx<double> q2_load(__amdgpu_buffer_rsrc_t r, unsigned off, double) {
  return __builtin_bit_cast(cplx<double>, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
__device__ __forceinline__ void q2_store(cplx<float> v, __amdgpu_buffer_rsrc_t r, unsigned off) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(q2u2, v), r, off, 0, 0);
}
__device__ __forceinline__ void q2_store(cplx<double> v, __amdgpu_buffer_rsrc_t r, unsigned off) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(q2u4, v), r, off, 0, 0);
}

// Z <- Q2 Z for a strip of 16 NWV columns of one matrix per workgroup; wave w
// owns columns 16 w .. 16 w + 15 of the strip.  The window of 2bw rows the
// current block acts on lives in LDS as two halves of bw rows (block k: slot
// k & 1 holds logical half 0, the other slot logical half 1), so consecutive
// blocks of a group share a half and each row of the strip is loaded and
// stored once per group.  Per block, on the matrix cores and for the wave's
// own columns only (no cross-wave dependency):
//   W = V^H Zw   (bw x 16, K = 2bw; reflectors i0.. meet rows i0 .. i0 + bw + 15)
//   Zw -= VT W   (2bw x 16, K = bw; VT(rr, i) = 0 for i <= rr - bw)
// with three real MFMAs per complex step in FP32 (P1 = Re a Re b, P2 = Im a
// Im b, P3 = (Re a +- Im a)(Re b + Im b), the operand sums formed in
// registers) and four in FP64.  The three-product form's imaginary part
// carries a cancellation that costs the FP32 vectors a small factor of
// accuracy; on the FP32 path the FP64 correction of the vectors that follows
// removes it.  V and VT of a block are shared by the waves: two barriers per
// block.
// The rows leave and enter the window while the matrix cores work: the next
// block's V and VT and the rows of its new half are loaded into registers
// during a block; at the start of the next block the half its predecessor
// finished is read from its slot, and its stores to HBM and the writes of the
// new rows into that slot are issued between the MFMAs of the first half of W
// (K over logical half 0, which is already in place), before the second half
// of W reads them.  Rows and columns outside the matrix go through the buffer
// unit's range check (see q2_load).
template <typename T, int BW, int NWV, bool ROWMAJOR>
__global__ void __launch_bounds__(WAVE * NWV) k_q2_apply(Q2Work<T> w) {
  constexpr bool M3 = sizeof(T) == 4;               // three real products per complex step
  constexpr int NT = WAVE * NWV;
  constexpr int SW = 16 * NWV;
  constexpr int VE = (2 * BW * BW + NT - 1) / NT;   // V (and VT) elements per thread
  constexpr int ZE = BW * 16 / WAVE;                // elements of the wave's Z half per lane
  constexpr int ES = (int)sizeof(cplx<T>);
  // rows padded by 128 bytes: the four rows one MFMA operand load spans then
  // fall in alternate halves of the LDS banks
  constexpr int PC = 128 / ES;
  __shared__ cplx<T> Zw[2][BW][SW + PC];
  __shared__ cplx<T> Vs[2 * BW][BW + PC];           // V(rr, i)
  __shared__ cplx<T> VTs[BW][2 * BW + PC];          // VT(rr, i) at VTs[i][rr]
  __shared__ cplx<T> Ws[BW][SW + PC];
  const int64_t strip = blockIdx.x;
  const int64_t n = w.n, b = blockIdx.y;
  const int tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
  const int cw = 16 * wave;                         // the wave's first column in the strip
  const int64_t col0 = strip * SW + cw;
  const cplx<T>* Vb = w.Vb + b * w.strideVTb;
  const cplx<T>* VTb = w.VTb + b * w.strideVTb;
  const int64_t G = ceil_div_dev(n - 1, BW);
  // the wave's columns of Z as a buffer resource (none when the wave lies
  // past the last column); element (row, c) at byte (row + c ldz) ES, or
  // (row ldz + c) ES for row-major Z (the host keeps the range below 2^31)
  const int64_t ncol = max<int64_t>(0, min<int64_t>(16, w.m - col0));
  const int64_t rbytes = ncol <= 0 ? 0
                         : ROWMAJOR ? ((n - 1) * w.ldz + ncol) * ES : ((ncol - 1) * w.ldz + n) * ES;
  const __amdgpu_buffer_rsrc_t zr = __builtin_amdgcn_make_buffer_rsrc(
      (void*)(w.Z + b * w.strideZ + (ncol > 0 ? (ROWMAJOR ? col0 : col0 * w.ldz) : 0)), (short)0,
      (int)rbytes, kRawBufferWord3);
  // Window moves, element q of a lane: column-major Z, lane -> row lane % bw
  // and column lane / bw + (64 / bw) q (each access reads whole 128-byte runs
  // of a column); row-major Z, lane -> column lane % 16 and row lane / 16 + 4 q
  // (16 lanes read the wave's 128 bytes of a row, the workgroup's waves the
  // strip's 1 KB; in LDS the four rows of an access fall in alternate bank
  // halves)
  const int zr_col = ROWMAJOR ? lane % 16 : lane / BW;
  const int zr_row = ROWMAJOR ? lane / 16 : lane % BW;
  auto zrow = [&](int q) { return ROWMAJOR ? zr_row + 4 * q : zr_row; };
  auto zcol = [&](int q) { return ROWMAJOR ? zr_col : zr_col + (WAVE / BW) * q; };
  auto zoff = [&](int64_t row0, int q, bool valid) -> unsigned {
    const int64_t row = row0 + zrow(q);
    const int64_t cc = zcol(q);
    if (!(valid && row < n && cc < ncol)) return kQ2Off;
    return (unsigned)((ROWMAJOR ? row * w.ldz + cc : row + cc * w.ldz) * ES);
  };
  auto zslot = [&](int slot, int q) -> cplx<T>& { return Zw[slot][zrow(q)][cw + zcol(q)]; };
  cplx<T> pv[VE], pvt[VE], pz[ZE], zst[ZE];
  auto fetch_v = [&](int64_t g, int64_t k) {
    const int64_t blk = (g * w.kmax + k) * (int64_t)(2 * BW * BW);
#pragma unroll
    for (int q = 0; q < VE; ++q) {
      const int e = tid + NT * q;
      const bool ok = e < 2 * BW * BW;
      pv[q] = ok ? Vb[blk + e] : cplx<T>{0, 0};
      pvt[q] = ok ? VTb[blk + e] : cplx<T>{0, 0};
    }
  };
  auto put_v = [&]() {
#pragma unroll
    for (int q = 0; q < VE; ++q) {
      const int e = tid + NT * q;
      if (e < 2 * BW * BW) {
        Vs[e / BW][e % BW] = pv[q];
        VTs[e / (2 * BW)][e % (2 * BW)] = pvt[q];
      }
    }
  };
  auto fetch_z = [&](int64_t row0) {
#pragma unroll
    for (int q = 0; q < ZE; ++q) pz[q] = q2_load(zr, zoff(row0, q, true), T(0));
  };
  auto wave_sync = []() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  for (int64_t g = G - 1; g >= 0; --g) {
    const int64_t j0 = g * BW;
    const int64_t Kg = hb2st_tasks(n, BW, j0);
    // logical half 0 of block 0 into slot 0; its half 1 waits in registers
    // for the first block's window move
    fetch_z(j0 + 1);
#pragma unroll
    for (int q = 0; q < ZE; ++q) zslot(0, q) = pz[q];
    fetch_z(j0 + 1 + BW);
    fetch_v(g, 0);
    for (int64_t k = 0; k < Kg; ++k) {
      const int64_t row0 = j0 + 1 + k * BW;
      const int s0 = (int)(k & 1), s1 = s0 ^ 1;      // slots of logical halves 0 and 1
      __syncthreads();                               // every wave is done with the previous V, VT
      put_v();
      __syncthreads();
      // the rows the previous block finished (slot s1; none before block 1)
#pragma unroll
      for (int q = 0; q < ZE; ++q) zst[q] = zslot(s1, q);
      cplx<T> pzn[ZE];
#pragma unroll
      for (int q = 0; q < ZE; ++q) pzn[q] = pz[q];
      if (k + 1 < Kg) {
        fetch_v(g, k + 1);
        fetch_z(row0 + 2 * BW);                      // the next block's new half
      }
      // W = V^H Zw on the wave's columns: all row tiles of W in one K loop
      // (their K ranges overlap and share the Zw operand), fully unrolled.
      // The LDS operands of step kk + 4 are read before the MFMAs of step kk
      // (scheduling barriers keep them there), so their latency hides behind
      // the matrix cores.  The first NMS steps also carry the window move
      // (stores of the finished rows, the new rows into slot s1); logical
      // half 1 is first read by the loads issued at step bw - 4, after all
      // of them.
      {
        using acc_t = typename mfma<T>::acc;
        constexpr int NRT = BW / 16;
        constexpr int KS = 2 * BW;
        constexpr int NMS = (BW / 4 - 1) < 4 ? (BW / 4 - 1) : 4;   // steps with a window move
        constexpr int MV = (ZE + NMS - 1) / NMS;     // window elements moved per such step
        // reflectors 16 t .. 16 t + 15 meet rows 16 t .. 16 t + bw + 15
        auto live = [](int t, int kk) { return kk >= 16 * t && kk < 16 * t + 16 + BW; };
        auto ld_z = [&](int kk) {
          const int kr = kk + lane / 16;
          return Zw[kr < BW ? s0 : s1][kr % BW][cw + lane % 16];
        };
        auto ld_v = [&](int kk, int t) { return Vs[kk + lane / 16][16 * t + lane % 16]; };
        acc_t p1[NRT], p2[NRT], p3[NRT];
#pragma unroll
        for (int t = 0; t < NRT; ++t) { p1[t] = acc_t{0, 0, 0, 0}; p2[t] = p1[t]; p3[t] = p1[t]; }
        cplx<T> bz = ld_z(0), v[NRT];
#pragma unroll
        for (int t = 0; t < NRT; ++t) v[t] = live(t, 0) ? ld_v(0, t) : cplx<T>{0, 0};
#pragma unroll
        for (int kk = 0; kk < KS; kk += 4) {
          cplx<T> bzn{0, 0}, vn[NRT];
#pragma unroll
          for (int t = 0; t < NRT; ++t) vn[t] = cplx<T>{0, 0};
          if (kk + 4 < KS) {
            bzn = ld_z(kk + 4);
#pragma unroll
            for (int t = 0; t < NRT; ++t)
              if (live(t, kk + 4)) vn[t] = ld_v(kk + 4, t);
          }
          __builtin_amdgcn_sched_barrier(0);
          if constexpr (M3) {
            // conj(v) z = P1 + P2 + i (P3 - P1 + P2), P1 = Re v Re z, P2 =
            // Im v Im z, P3 = (Re v - Im v)(Re z + Im z)
            const T zs = bz.x + bz.y;
#pragma unroll
            for (int t = 0; t < NRT; ++t) {
              if (!live(t, kk)) continue;
              p1[t] = mfma<T>::op(v[t].x, bz.x, p1[t]);
              p2[t] = mfma<T>::op(v[t].y, bz.y, p2[t]);
              p3[t] = mfma<T>::op(v[t].x - v[t].y, zs, p3[t]);
            }
          } else {
#pragma unroll
            for (int t = 0; t < NRT; ++t) {
              // conj(v) z with four real products (p1 = Re, p2 = Im)
              if (!live(t, kk)) continue;
              p1[t] = mfma<T>::op(v[t].x, bz.x, p1[t]);
              p1[t] = mfma<T>::op(v[t].y, bz.y, p1[t]);
              p2[t] = mfma<T>::op(v[t].x, bz.y, p2[t]);
              p2[t] = mfma<T>::op(-v[t].y, bz.x, p2[t]);
            }
          }
          if (kk < 4 * NMS) {
#pragma unroll
            for (int u = 0; u < MV; ++u) {
              const int q = (kk / 4) * MV + u;
              if (q >= ZE) continue;
              q2_store(zst[q], zr, zoff(row0 - BW, q, k > 0));
              zslot(s1, q) = pzn[q];
            }
          }
          __builtin_amdgcn_sched_barrier(0);
          bz = bzn;
#pragma unroll
          for (int t = 0; t < NRT; ++t) v[t] = vn[t];
        }
#pragma unroll
        for (int t = 0; t < NRT; ++t)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            Ws[16 * t + mfma<T>::row(lane, r)][cw + lane % 16] =
                M3 ? cplx<T>{p1[t][r] + p2[t][r], p3[t][r] - p1[t][r] + p2[t][r]}
                   : cplx<T>{p1[t][r], p2[t][r]};
      }
      wave_sync();
      // Zw -= VT W on the wave's columns: the row tiles share the W operand;
      // operands of the next step read ahead as above
      {
        using acc_t = typename mfma<T>::acc;
        constexpr int NRT = 2 * BW / 16;
        // VT(rr, i) = 0 for i <= rr - bw
        auto live = [](int t, int kk) { return kk + 3 >= 16 * t - BW + 1; };
        auto ld_w = [&](int kk) { return Ws[kk + lane / 16][cw + lane % 16]; };
        auto ld_a = [&](int kk, int t) { return VTs[kk + lane / 16][16 * t + lane % 16]; };
        acc_t p1[NRT], p2[NRT], p3[NRT];
#pragma unroll
        for (int t = 0; t < NRT; ++t) { p1[t] = acc_t{0, 0, 0, 0}; p2[t] = p1[t]; p3[t] = p1[t]; }
        cplx<T> wv = ld_w(0), a[NRT];
#pragma unroll
        for (int t = 0; t < NRT; ++t) a[t] = live(t, 0) ? ld_a(0, t) : cplx<T>{0, 0};
#pragma unroll
        for (int kk = 0; kk < BW; kk += 4) {
          cplx<T> wn{0, 0}, an[NRT];
#pragma unroll
          for (int t = 0; t < NRT; ++t) an[t] = cplx<T>{0, 0};
          if (kk + 4 < BW) {
            wn = ld_w(kk + 4);
#pragma unroll
            for (int t = 0; t < NRT; ++t)
              if (live(t, kk + 4)) an[t] = ld_a(kk + 4, t);
          }
          __builtin_amdgcn_sched_barrier(0);
          if constexpr (M3) {
            // a w = P1 - P2 + i (P3 - P1 - P2)
            const T ws = wv.x + wv.y;
#pragma unroll
            for (int t = 0; t < NRT; ++t) {
              if (!live(t, kk)) continue;
              p1[t] = mfma<T>::op(a[t].x, wv.x, p1[t]);
              p2[t] = mfma<T>::op(a[t].y, wv.y, p2[t]);
              p3[t] = mfma<T>::op(a[t].x + a[t].y, ws, p3[t]);
            }
          } else {
#pragma unroll
            for (int t = 0; t < NRT; ++t) {
              if (!live(t, kk)) continue;
              p1[t] = mfma<T>::op(a[t].x, wv.x, p1[t]);
              p1[t] = mfma<T>::op(-a[t].y, wv.y, p1[t]);
              p2[t] = mfma<T>::op(a[t].x, wv.y, p2[t]);
              p2[t] = mfma<T>::op(a[t].y, wv.x, p2[t]);
            }
          }
          __builtin_amdgcn_sched_barrier(0);
          wv = wn;
#pragma unroll
          for (int t = 0; t < NRT; ++t) a[t] = an[t];
        }
#pragma unroll
        for (int t = 0; t < NRT; ++t)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int rr = 16 * t + mfma<T>::row(lane, r);
            cplx<T>& zz = Zw[rr < BW ? s0 : s1][rr % BW][cw + lane % 16];
            zz = csub(zz, M3 ? cplx<T>{p1[t][r] - p2[t][r], p3[t][r] - p1[t][r] - p2[t][r]}
                             : cplx<T>{p1[t][r], p2[t][r]});
          }
      }
      wave_sync();
    }
    // the group's last window: both halves are finished
    const int64_t rowL = j0 + 1 + (Kg - 1) * BW;
    const int sL = (int)((Kg - 1) & 1);
#pragma unroll
    for (int q = 0; q < ZE; ++q) {
      q2_store(zslot(sL, q), zr, zoff(rowL, q, true));
      q2_store(zslot(sL ^ 1, q), zr, zoff(rowL + BW, q, true));
    }
    wave_sync();
  }
}

template <typename T>
void prepare_q2(const Q2Work<T>& w, hipStream_t s) {
  const int64_t n = w.n, B = w.batch;
  if (n < 2 || w.m <= 0) return;
  const int64_t G = ceil_div(n - 1, w.bw);
  const dim3 grid((unsigned)w.kmax, (unsigned)G, (unsigned)B);
  if (w.bw == 32) hipLaunchKernelGGL((k_q2_prep<T, 32>), grid, dim3(256), 0, s, w);
  else if (w.bw == 16) hipLaunchKernelGGL((k_q2_prep<T, 16>), grid, dim3(256), 0, s, w);
  else throw std::runtime_error("prepare_q2: band width must be 16 or 32");
  PPH_CHECK_LAUNCH();
}

template <typename T>
void apply_q2(const Q2Work<T>& w, hipStream_t s) {
  const int64_t n = w.n, B = w.batch;
  if (n < 2 || w.m <= 0) return;
  // columns per workgroup: 16 per wave, 8 waves for FP32 and 2 for FP64 (LDS)
  constexpr int NWV = sizeof(T) == 4 ? 8 : 2;
  const dim3 grid((unsigned)ceil_div(w.m, 16 * NWV), (unsigned)B);
  if (w.bw != 32 && w.bw != 16) throw std::runtime_error("apply_q2: band width must be 16 or 32");
  if (w.rowmajor) {
    if (w.ldz < w.m || n * w.ldz * (int64_t)sizeof(cplx<T>) > 0x7FFFFFFF)
      throw std::runtime_error("apply_q2: row-major Z needs m <= ldz and n ldz elements below 2^31 bytes");
    if (w.bw == 32) hipLaunchKernelGGL((k_q2_apply<T, 32, NWV, true>), grid, dim3(WAVE * NWV), 0, s, w);
    else hipLaunchKernelGGL((k_q2_apply<T, 16, NWV, true>), grid, dim3(WAVE * NWV), 0, s, w);
  } else {
    if ((16 * w.ldz) * (int64_t)sizeof(cplx<T>) > 0x7FFFFFFF)
      throw std::runtime_error("apply_q2: 16 columns of Z must stay below 2^31 bytes");
    if (w.bw == 32) hipLaunchKernelGGL((k_q2_apply<T, 32, NWV, false>), grid, dim3(WAVE * NWV), 0, s, w);
    else hipLaunchKernelGGL((k_q2_apply<T, 16, NWV, false>), grid, dim3(WAVE * NWV), 0, s, w);
  }
  PPH_CHECK_LAUNCH();
}

template void prepare_q2<float>(const Q2Work<float>&, hipStream_t);
template void prepare_q2<double>(const Q2Work<double>&, hipStream_t);
template void apply_q2<float>(const Q2Work<float>&, hipStream_t);
template void apply_q2<double>(const Q2Work<double>&, hipStream_t);

}  // namespace pph
