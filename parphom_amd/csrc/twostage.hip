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
//          as GEMMs on the matrix cores.  The trailing matrix is read three
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

// ---------------------------------------------------------------- Hermitian fill
// A(r, c) for r < c := conj(A(c, r)) and Im A(c, c) := 0, from the lower
// triangle (the triangle the library's reductions treat as authoritative).
template <typename T>
__global__ void __launch_bounds__(256) k_herm_from_lower(cplx<T>* A, int64_t lda, int64_t strideA,
                                                         int64_t n) {
  __shared__ cplx<T> t[32][33];
  const int64_t ti = blockIdx.y, tj = blockIdx.x;      // source tile: rows ti, columns tj
  if (tj > ti) return;
  cplx<T>* M = A + blockIdx.z * strideA;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
  for (int r = ty; r < 32; r += 8) {
    const int64_t row = ti * 32 + tx, col = tj * 32 + r;
    t[tx][r] = (row < n && col < n) ? M[row + col * lda] : cplx<T>{0, 0};
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    // destination (tj * 32 + tx, ti * 32 + r) = conj(source (ti * 32 + r, tj * 32 + tx))
    const int64_t drow = tj * 32 + tx, dcol = ti * 32 + r;
    if (drow >= n || dcol >= n) continue;
    const cplx<T> v = t[r][tx];
    if (dcol > drow) M[drow + dcol * lda] = cplx<T>{v.x, -v.y};
    else if (dcol == drow) M[drow + dcol * lda] = cplx<T>{v.x, T(0)};
  }
}

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
      double zr = 0.0, zi = 0.0;
      for (int64_t r0_ = i; r0_ < m; r0_ += 4 * WAVE) {
        cplx<T> vv[4], xx[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int64_t r = r0_ + lane + u * WAVE;
          const bool ok = r < m;
          vv[u] = ok ? Vi[r] : cplx<T>{0, 0};
          xx[u] = ok ? Pc[r] : cplx<T>{0, 0};
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          // conj(v) x
          zr += (double)vv[u].x * xx[u].x + (double)vv[u].y * xx[u].y;
          zi += (double)vv[u].x * xx[u].y - (double)vv[u].y * xx[u].x;
        }
      }
      zr = wave_sum(zr);
      zi = wave_sum(zi);
      // P(:, c) -= conj(tau) z v
      const cplx<T> f{(T)(ctr * zr - cti * zi), (T)(ctr * zi + cti * zr)};
      const bool next = (c == i + 1);
      double s = 0.0;
      for (int64_t r0_ = i; r0_ < m; r0_ += 4 * WAVE) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int64_t r = r0_ + lane + u * WAVE;
          if (r >= m) continue;
          const cplx<T> x = csub(Pc[r], cmul(f, Vi[r]));
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
  {
    const int64_t nt = ceil_div(n, 32);
    hipLaunchKernelGGL(k_herm_from_lower<T>, dim3((unsigned)nt, (unsigned)nt, (unsigned)B),
                       dim3(256), 0, s, w.A, w.lda, w.strideA, n);
    PPH_CHECK_LAUNCH();
  }
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
      launch_gemm<T>(g, s);
    }
    // X = A22 VT   (m x bw)
    {
      GemmParams<T> g{};
      g.opA = 'N'; g.opB = 'N';
      g.m = m; g.n = bw; g.k = m;
      g.alpha = cplx<T>{1, 0}; g.beta = cplx<T>{0, 0};
      g.A = w.A + r0 + r0 * w.lda; g.lda = w.lda; g.strideA = w.strideA;
      g.B = w.VT + r0; g.ldb = n; g.strideB = w.strideXV;
      g.C = w.X + r0; g.ldc = n; g.strideC = w.strideXV;
      g.batch = B;
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
      launch_gemm<T>(g, s);
    }
    // A22 -= Y V^H + V Y^H   (both triangles, so the next panel's products
    // can read A22 whole)
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
//   which is the next step's X (kept in registers).
// A wave holds a bw x bw block as lane = row r (lane % bw) and column group
// lane / bw of bw^2 / 64 consecutive columns.  Band element (r, c), r >= c,
// sits at AB[c * 2bw + r - c].
template <typename T, int BW>
struct HbLane {
  static constexpr int NG = WAVE / BW;     // column groups
  static constexpr int CPL = BW / NG;      // columns per lane
};

template <typename T, int BW>
__device__ __forceinline__ T sum_rows(T v) {        // over the bw lanes of a column group
#pragma unroll
  for (int o = 1; o < BW; o <<= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}
template <typename T, int BW>
__device__ __forceinline__ T sum_groups(T v) {      // over the column groups of a row
#pragma unroll
  for (int o = BW; o < WAVE; o <<= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

template <typename T, int BW, int NW>
__global__ void __launch_bounds__(WAVE * NW) k_hb2st(TwoStageWork<T> w) {
  constexpr int CPL = HbLane<T, BW>::CPL;
  constexpr int LDB = 2 * BW;
  constexpr int SH = 4096;                 // progress = sweep * SH + steps done
  const int64_t b = blockIdx.x;
  const int64_t n = w.n;
  cplx<T>* AB = w.AB + b * w.strideAB;
  cplx<T>* V2 = w.V2 + b * w.strideV2;
  cplx<T>* tau2 = w.tau2 + b * w.strideTau2;
  const int64_t kmax = w.kmax;
  const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
  const int r = lane % BW, c0 = (lane / BW) * CPL;
  __shared__ int prog[NW];
  __shared__ cplx<T> sv[NW][BW], sw[NW][BW];
  __shared__ int stalled;
  if (threadIdx.x < NW) prog[threadIdx.x] = -1;
  if (threadIdx.x == 0) stalled = 0;
  __syncthreads();
  const int pw = (wave + NW - 1) % NW;
  for (int64_t j = wave; j < n - 1; j += NW) {
    const int64_t K = hb2st_tasks(n, BW, j);
    const int64_t Kp = (j > 0) ? hb2st_tasks(n, BW, j - 1) : 0;
    cplx<T> X[CPL];
#pragma unroll
    for (int i = 0; i < CPL; ++i) X[i] = cplx<T>{0, 0};
    for (int64_t k = 0; k < K; ++k) {
      if (j > 0) {
        // sweep j - 1 must have finished its steps 0 .. k + 2.  The wait is
        // bounded (about 0.2 s, far beyond any step): a stall is flagged in
        // info and every later wait of the workgroup is skipped, so the
        // kernel always ends
        const int need = (int)((j - 1) * SH + min<int64_t>(k + 3, Kp));
        int spins = 0;
        while (__hip_atomic_load(&prog[pw], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < need &&
               __hip_atomic_load(&stalled, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0) {
          __builtin_amdgcn_s_sleep(2);
          if (++spins > (1 << 22)) {
            if (lane == 0)
              __hip_atomic_store(&stalled, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            break;
          }
        }
      }
      const int64_t s = j + 1 + k * BW;
      const int L = (int)min<int64_t>(BW, n - s);
      const int64_t sp = (k == 0) ? j : s - BW;      // first column of X
      if (k == 0) {
        if (c0 == 0 && r < L) X[0] = AB[j * LDB + (s + r - j)];
      }
      // reflector from the first column of X (lanes of column group 0, slot 0)
      double xr = 0.0, xi = 0.0, nn = 0.0;
      if (c0 == 0 && r < L) {
        xr = (double)X[0].x;
        xi = (double)X[0].y;
        if (r >= 1) nn = xr * xr + xi * xi;
      }
      nn = wave_sum(nn);
      const double ar = __shfl(xr, 0, WAVE), ai = __shfl(xi, 0, WAVE);
      const Reflector rf = make_reflector(nn, ar, ai);
      const double xrr = __shfl(xr, r, WAVE), xir = __shfl(xi, r, WAVE);
      cplx<T> v{0, 0};
      if (r == 0) v = cplx<T>{1, 0};
      else if (r < L) v = cplx<T>{(T)(xrr * rf.sr - xir * rf.si), (T)(xrr * rf.si + xir * rf.sr)};
      const cplx<T> tau{(T)rf.tr, (T)rf.ti};
      if (c0 == 0) sv[wave][r] = v;
      if (k >= 1) {
        // X <- H^H X on columns 1 ..: z_c = sum_r conj(v_r) X(r, c)
        const cplx<T> ctau{tau.x, -tau.y};
#pragma unroll
        for (int i = 0; i < CPL; ++i) {
          T zr = v.x * X[i].x + v.y * X[i].y;
          T zi = v.x * X[i].y - v.y * X[i].x;
          zr = sum_rows<T, BW>(zr);
          zi = sum_rows<T, BW>(zi);
          if (c0 + i >= 1) X[i] = csub(X[i], cmul(cmul(ctau, v), cplx<T>{zr, zi}));
        }
      }
      if (c0 == 0) X[0] = (r == 0) ? cplx<T>{(T)rf.beta, 0} : cplx<T>{0, 0};
      if (k == 0) {
        if (c0 == 0 && r < L) AB[j * LDB + (s + r - j)] = X[0];
      } else if (r < L) {
#pragma unroll
        for (int i = 0; i < CPL; ++i) {
          const int64_t col = sp + c0 + i;
          AB[col * LDB + (s + r - col)] = X[i];
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      cplx<T> vc[CPL];
#pragma unroll
      for (int i = 0; i < CPL; ++i) vc[i] = (c0 + i < L) ? sv[wave][c0 + i] : cplx<T>{0, 0};
      // D <- H^H D H on the diagonal block (Hermitian; lower triangle stored)
      cplx<T> D[CPL];
#pragma unroll
      for (int i = 0; i < CPL; ++i) {
        const int c = c0 + i;
        cplx<T> d{0, 0};
        if (r < L && c < L) {
          if (r >= c) d = AB[(s + c) * LDB + (r - c)];
          else {
            const cplx<T> u = AB[(s + r) * LDB + (c - r)];
            d = cplx<T>{u.x, -u.y};
          }
        }
        D[i] = d;
      }
      cplx<T> y{0, 0};
#pragma unroll
      for (int i = 0; i < CPL; ++i) y = cadd(y, cmul(D[i], vc[i]));
      y.x = sum_groups<T, BW>(y.x);
      y.y = sum_groups<T, BW>(y.y);
      double h = (c0 == 0 && r < L) ? (double)v.x * y.x + (double)v.y * y.y : 0.0;
      h = wave_sum(h);                                // v^H D v (real)
      const double t2 = rf.tr * rf.tr + rf.ti * rf.ti;
      const cplx<T> ty = cmul(tau, y);
      const T hh = (T)(0.5 * t2 * h);
      const cplx<T> wr{ty.x - hh * v.x, ty.y - hh * v.y};
      if (c0 == 0) sw[wave][r] = wr;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int i = 0; i < CPL; ++i) {
        const int c = c0 + i;
        if (r < L && c < L && r >= c) {
          const cplx<T> wc = sw[wave][c];
          // D -= w v^H + v w^H
          cplx<T> d = csub(D[i], cmulc(vc[i], wr));            // w_r conj(v_c)
          d = csub(d, cmulc(wc, v));                           // v_r conj(w_c)
          if (r == c) d.y = T(0);
          AB[(s + c) * LDB + (r - c)] = d;
        }
      }
      // Y <- Y H on the block below, carried as the next step's X
      if (k + 1 < K) {
        const int64_t s2 = s + BW;
        const int L2 = (int)min<int64_t>(BW, n - s2);
        cplx<T> t{0, 0};
#pragma unroll
        for (int i = 0; i < CPL; ++i) {
          const int c = c0 + i;
          cplx<T> yv{0, 0};
          if (r < L2 && c < L) yv = AB[(s + c) * LDB + (s2 + r - s - c)];
          X[i] = yv;
          t = cadd(t, cmul(yv, vc[i]));
        }
        t.x = sum_groups<T, BW>(t.x);
        t.y = sum_groups<T, BW>(t.y);
        const cplx<T> tt = cmul(tau, t);
#pragma unroll
        for (int i = 0; i < CPL; ++i) {
          const int c = c0 + i;
          X[i] = csub(X[i], cmulc(vc[i], tt));                   // tau t_r conj(v_c)
          if (r < L2 && c < L) AB[(s + c) * LDB + (s2 + r - s - c)] = X[i];
        }
      }
      // the reflector for the back-transformation
      if (c0 == 0) V2[(j * kmax + k) * BW + r] = v;
      if (lane == 0) tau2[j * kmax + k] = tau;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (lane == 0)
        __hip_atomic_store(&prog[wave], (int)(j * SH + k + 1), __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
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

// VT = V T per block (column-major 2bw x bw at VTb[(g kmax + k) 2bw^2]).
// One workgroup per (k, g, matrix).
template <typename T, int BW>
__global__ void __launch_bounds__(256) k_q2_prep(Q2Work<T> w) {
  __shared__ cplx<T> Vs[2 * BW][BW + 1];
  __shared__ cplx<T> ts[BW];
  __shared__ cplx<double> G[BW][BW + 1], Tm[BW][BW + 1];
  const int64_t k = blockIdx.x, g = blockIdx.y, b = blockIdx.z;
  const int64_t n = w.n, j0 = g * BW;
  if (j0 >= n - 1 || k >= hb2st_tasks(n, BW, j0)) return;
  const int tid = threadIdx.x;
  gather_v<T, BW>(w.V2 + b * w.strideV2, w.tau2 + b * w.strideTau2, w.kmax, n, j0, k, Vs, ts);
  __syncthreads();
  // G = V^H V (column i nonzero on rows i .. i + bw - 1)
  for (int e = tid; e < BW * BW; e += 256) {
    const int a = e % BW, c = e / BW;
    const int lo = max(a, c), hi = min(a, c) + BW;
    double sr = 0.0, si = 0.0;
    for (int rr = lo; rr < hi; ++rr) {
      const cplx<T> x = Vs[rr][a], y = Vs[rr][c];
      sr += (double)x.x * y.x + (double)x.y * y.y;
      si += (double)x.x * y.y - (double)x.y * y.x;
    }
    G[a][c] = cplx<double>{sr, si};
    Tm[a][c] = cplx<double>{0, 0};
  }
  __syncthreads();
  for (int i = 0; i < BW; ++i) {
    const cplx<double> ti{(double)ts[i].x, (double)ts[i].y};
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
  // VT(rr, c) = sum_{l <= c} V(rr, l) T(l, c)
  cplx<T>* out = w.VTb + b * w.strideVTb + (g * w.kmax + k) * (int64_t)(2 * BW * BW);
  for (int e = tid; e < 2 * BW * BW; e += 256) {
    const int rr = e % (2 * BW), c = e / (2 * BW);
    double sr = 0.0, si = 0.0;
    for (int l = max(0, rr - BW + 1); l <= min(c, rr); ++l) {
      const cplx<T> x = Vs[rr][l];
      const cplx<double> t = Tm[l][c];
      sr += (double)x.x * t.x - (double)x.y * t.y;
      si += (double)x.x * t.y + (double)x.y * t.x;
    }
    out[rr + c * (2 * BW)] = cplx<T>{(T)sr, (T)si};
  }
}

// Z <- Q2 Z for a strip of SW columns of one matrix per workgroup; the window
// of 2bw rows the current block acts on lives in LDS as two halves of bw rows
// (physical half (k + h) & 1 holds logical half h), so consecutive blocks of
// a group share one half and each row of the strip is loaded and stored once
// per group.
template <typename T, int BW, int SW>
__global__ void __launch_bounds__(256) k_q2_apply(Q2Work<T> w) {
  constexpr int NT = 256;
  constexpr int IW = BW * SW / NT;          // reflectors per thread in W = V^H Zw
  constexpr int RW = 2 * BW * SW / NT;      // rows per thread in Zw -= VT W
  static_assert(NT % SW == 0 && IW >= 1 && RW >= 1, "strip shape");
  __shared__ cplx<T> Zw[2][BW][SW + 1];
  __shared__ cplx<T> Vs[2 * BW][BW + 1];
  __shared__ cplx<T> VTs[BW][2 * BW + 1];   // VTs[i][rr] = VT(rr, i)
  __shared__ cplx<T> Ws[BW][SW + 1];
  __shared__ cplx<T> ts[BW];
  const int64_t n = w.n, b = blockIdx.y;
  const int64_t col0 = (int64_t)blockIdx.x * SW;
  const int tid = threadIdx.x;
  const int c = tid % SW;
  const int grp = tid / SW;
  cplx<T>* Z = w.Z + b * w.strideZ;
  const cplx<T>* V2 = w.V2 + b * w.strideV2;
  const cplx<T>* tau2 = w.tau2 + b * w.strideTau2;
  const cplx<T>* VTb = w.VTb + b * w.strideVTb;
  const int64_t G = ceil_div_dev(n - 1, BW);
  // half h of the window at block k: rows j0 + 1 + (k + h) bw .. , slot (k + h) & 1
  auto load_half = [&](int64_t row0, int slot) {
    for (int e = tid; e < BW * SW; e += NT) {
      const int rr = e % BW, cc = e / BW;
      const int64_t row = row0 + rr, col = col0 + cc;
      Zw[slot][rr][cc] = (row < n && col < w.m) ? Z[row + col * w.ldz] : cplx<T>{0, 0};
    }
  };
  auto store_half = [&](int64_t row0, int slot) {
    for (int e = tid; e < BW * SW; e += NT) {
      const int rr = e % BW, cc = e / BW;
      const int64_t row = row0 + rr, col = col0 + cc;
      if (row < n && col < w.m) Z[row + col * w.ldz] = Zw[slot][rr][cc];
    }
  };
  for (int64_t g = G - 1; g >= 0; --g) {
    const int64_t j0 = g * BW;
    const int64_t Kg = hb2st_tasks(n, BW, j0);
    __syncthreads();
    load_half(j0 + 1, 0);
    load_half(j0 + 1 + BW, 1);
    for (int64_t k = 0; k < Kg; ++k) {
      const int64_t row0 = j0 + 1 + k * BW;
      gather_v<T, BW>(V2, tau2, w.kmax, n, j0, k, Vs, ts);
      const cplx<T>* vt = VTb + (g * w.kmax + k) * (int64_t)(2 * BW * BW);
      for (int e = tid; e < 2 * BW * BW; e += NT) {
        const int rr = e % (2 * BW), i = e / (2 * BW);
        VTs[i][rr] = vt[rr + i * (2 * BW)];
      }
      __syncthreads();
      const int s0 = (int)(k & 1);         // slot of logical half 0
      // W(i, c) = sum_rr conj(V(rr, i)) Zw(rr, c), i = grp IW .. + IW
      {
        const int i0 = grp * IW;
        cplx<T> acc[IW];
#pragma unroll
        for (int q = 0; q < IW; ++q) acc[q] = cplx<T>{0, 0};
        const int rlo = i0, rhi = min(2 * BW, i0 + IW - 1 + BW);
        for (int rr = rlo; rr < rhi; ++rr) {
          const cplx<T> z = Zw[(s0 + rr / BW) & 1][rr % BW][c];
#pragma unroll
          for (int q = 0; q < IW; ++q) {
            const cplx<T> v = Vs[rr][i0 + q];
            acc[q].x += v.x * z.x + v.y * z.y;
            acc[q].y += v.x * z.y - v.y * z.x;
          }
        }
#pragma unroll
        for (int q = 0; q < IW; ++q) Ws[i0 + q][c] = acc[q];
      }
      __syncthreads();
      // Zw(rr, c) -= sum_i VT(rr, i) W(i, c), rows grp RW .. + RW; VT(rr, i) = 0
      // for rr >= i + bw
      {
        const int rr0 = grp * RW;
        cplx<T> acc[RW];
#pragma unroll
        for (int q = 0; q < RW; ++q) acc[q] = cplx<T>{0, 0};
        for (int i = max(0, rr0 - BW + 1); i < BW; ++i) {
          const cplx<T> wv = Ws[i][c];
#pragma unroll
          for (int q = 0; q < RW; ++q) {
            const cplx<T> a = VTs[i][rr0 + q];
            acc[q].x += a.x * wv.x - a.y * wv.y;
            acc[q].y += a.x * wv.y + a.y * wv.x;
          }
        }
#pragma unroll
        for (int q = 0; q < RW; ++q) {
          const int rr = rr0 + q;
          cplx<T>& zz = Zw[(s0 + rr / BW) & 1][rr % BW][c];
          zz = csub(zz, acc[q]);
        }
      }
      __syncthreads();
      // logical half 0 is finished for this group: store it, and bring in the
      // half after the window into its slot
      store_half(row0, s0);
      if (k + 1 < Kg) {
        __syncthreads();
        load_half(row0 + 2 * BW, s0);
      } else {
        store_half(row0 + BW, s0 ^ 1);
      }
      __syncthreads();
    }
  }
}

template <typename T>
void apply_q2(const Q2Work<T>& w, hipStream_t s) {
  const int64_t n = w.n, B = w.batch;
  if (n < 2 || w.m <= 0) return;
  const int64_t G = ceil_div(n - 1, w.bw);
  if (w.bw == 32) {
    hipLaunchKernelGGL((k_q2_prep<T, 32>), dim3((unsigned)w.kmax, (unsigned)G, (unsigned)B),
                       dim3(256), 0, s, w);
    PPH_CHECK_LAUNCH();
    constexpr int SW = sizeof(T) == 4 ? 64 : 32;
    hipLaunchKernelGGL((k_q2_apply<T, 32, SW>), dim3((unsigned)ceil_div(w.m, SW), (unsigned)B),
                       dim3(256), 0, s, w);
    PPH_CHECK_LAUNCH();
  } else if (w.bw == 16) {
    hipLaunchKernelGGL((k_q2_prep<T, 16>), dim3((unsigned)w.kmax, (unsigned)G, (unsigned)B),
                       dim3(256), 0, s, w);
    PPH_CHECK_LAUNCH();
    hipLaunchKernelGGL((k_q2_apply<T, 16, 64>), dim3((unsigned)ceil_div(w.m, 64), (unsigned)B),
                       dim3(256), 0, s, w);
    PPH_CHECK_LAUNCH();
  } else {
    throw std::runtime_error("apply_q2: band width must be 16 or 32");
  }
}

template void apply_q2<float>(const Q2Work<float>&, hipStream_t);
template void apply_q2<double>(const Q2Work<double>&, hipStream_t);

}  // namespace pph
