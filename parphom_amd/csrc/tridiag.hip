// Hermitian -> real symmetric tridiagonal reduction, T = Q^H A Q, batched.
//
// Blocked Householder reduction on the lower triangle: panels of `nb`
// reflectors are formed column by column (the LATRD scheme: the trailing
// matrix is not touched inside a panel; each new column is corrected with the
// panel's V and W), then the trailing matrix gets one rank-2nb update
// A -= V W^H + W V^H on the matrix cores.  Per column three kernels run:
//   k_col   : alpha2 of the previous column (from k_fin's partial dots),
//             correct column j with the panel so far, per-workgroup partial
//             sums of |x|^2.
//   k_hemv  : u = S x for the trailing matrix S = A(j+1:, j+1:) and the
//             unscaled column x = (0, A(j+2:, j)) over a 2-D grid of tile rows
//             x column segments, stored as per-workgroup partials (row sums
//             per segment, column sums per tile row), plus the dots W^H x,
//             V^H x per tile row; workgroup (0, 0) forms the reflector (beta,
//             tau, scale = 1/(alpha-beta)) from k_col's partials.
//   k_fin   : sums the hemv's partials in a fixed order, then with
//             v = e_0 + scale x: y = S(:, 0) + scale u, the dots
//             W^H v = conj(W(j+1, :)) + scale W^H x (same for V), then
//             y - V (W^H v) - W (V^H v), w = tau y; writes v into V and
//             per-workgroup partials of w^H v.
// Keeping the hemv linear in the unscaled x lets it start without waiting for
// the reflector.
// Every cross-workgroup reduction ends at a kernel boundary: producers store
// per-workgroup partials and the next kernel sums them in a fixed order, so no
// kernel needs device-scope fences or atomics, and the results repeat bit for
// bit from run to run.
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
constexpr int S_TAU = 3, S_SCALE = 5, S_A2 = 8;

// Sum of np per-workgroup partials (stride `st` doubles apart), in a fixed order.
__device__ __forceinline__ double sum_partials(const double* x, int64_t np, int st, double* sh) {
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < np; i += blockDim.x) s += x[i * st];
  return block_sum(s, sh);
}


template <typename T>
__device__ __forceinline__ cplx<T> weff_a(const TridiagWork<T>& w, int64_t b, int64_t r, int k,
                                          int64_t p, cplx<T> a2) {
  const cplx<T> wv = w.W[b * w.strideW + r + (int64_t)k * w.n];
  const cplx<T> vv = w.V[b * w.strideV + r + (p + k) * w.n];
  return cadd(wv, cmul(a2, vv));
}

template <typename T>
__device__ __forceinline__ cplx<T> weff(const TridiagWork<T>& w, int64_t b, int64_t r, int k,
                                        int64_t p, const double* scal) {
  return weff_a(w, b, r, k, p, cplx<T>{(T)scal[S_A2 + 2 * k], (T)scal[S_A2 + 2 * k + 1]});
}

// alpha2 = -tau/2 (w^H v) for the column whose k_fin ran over m rows: its
// per-workgroup partials of w^H v sit at part[yv + 2 i], tau at scal[S_TAU].
__device__ __forceinline__ void alpha2_from_partials(const double* part, int64_t yv, int64_t m,
                                                     const double* scal, double* sh, double& a2r,
                                                     double& a2i) {
  const int64_t np = (m + 63) / 64;
  const double yr = sum_partials(part + yv, np, 2, sh);
  const double yi = sum_partials(part + yv + 1, np, 2, sh);
  const double tr = scal[S_TAU], ti = scal[S_TAU + 1];
  a2r = -0.5 * (tr * yr - ti * yi);
  a2i = -0.5 * (tr * yi + ti * yr);
}

// Row r of the panel terms sum_k X(r, k) a_k + Y(r, k) b_k for k = sub, sub + 4,
// ..., < jj: the loads of eight terms are issued together from clamped
// addresses and masked afterwards, so a wave waits for one memory latency per
// eight terms instead of one per term.
template <typename T, typename FX, typename FY>
__device__ __forceinline__ cplx<T> panel_terms(int sub, int jj, FX&& term_x, FY&& term_y) {
  cplx<T> acc{0, 0};
  for (int k0 = sub; k0 < jj; k0 += 32) {
    cplx<T> tx[8], ty[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = k0 + 4 * i;
      const bool ok = k < jj;
      const int kc = ok ? k : sub;
      tx[i] = term_x(kc);
      ty[i] = term_y(kc);
      if (!ok) { tx[i] = cplx<T>{0, 0}; ty[i] = cplx<T>{0, 0}; }
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) acc = cadd(acc, cadd(tx[i], ty[i]));
  }
  return acc;
}

// Raw V(r, p + k) and W(r, k) of a thread's first eight panel terms,
// k = sub + 4 i (i < 8), loaded before the kernel's first barrier so that
// their latency overlaps its reductions (a barrier keeps the compiler from
// moving loads above it).  Entries with k >= jj or !ok are loaded from a
// valid address and never used; terms from k = sub + 32 on (panels wider
// than 32) are loaded where they are used.
template <typename T>
struct PanelRow {
  cplx<T> v[8], w[8];
};
template <typename T>
__device__ __forceinline__ void load_panel_row(const TridiagWork<T>& w, int64_t b, int64_t r,
                                               int sub, int jj, int64_t p, bool ok,
                                               PanelRow<T>& pr) {
  const int64_t rc = ok ? r : 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int k = sub + 4 * i;
    const int kc = (ok && k < jj) ? k : 0;
    pr.v[i] = w.V[b * w.strideV + rc + (p + kc) * w.n];
    pr.w[i] = w.W[b * w.strideW + rc + (int64_t)kc * w.n];
  }
}
// sum over the thread's terms of V(r, k) x_k + W_eff(r, k) y_k, with
// W_eff(r, k) = W(r, k) + a2(k) V(r, k): the first eight from pr, the rest
// (k >= sub + 32) loaded here
template <typename T, typename FA, typename FX, typename FY>
__device__ __forceinline__ cplx<T> panel_row_terms(const TridiagWork<T>& w, int64_t b, int64_t r,
                                                   int sub, int jj, int64_t p,
                                                   const PanelRow<T>& pr, FA&& a2, FX&& xk,
                                                   FY&& yk) {
  cplx<T> acc{0, 0};
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int k = sub + 4 * i;
    if (k < jj) {
      const cplx<T> we = cadd(pr.w[i], cmul(a2(k), pr.v[i]));
      acc = cadd(acc, cadd(cmul(pr.v[i], xk(k)), cmul(we, yk(k))));
    }
  }
  if (sub + 32 < jj) {
    acc = cadd(acc, panel_terms<T>(
        sub + 32, jj, [&](int k) { return cmul(w.V[b * w.strideV + r + (p + k) * w.n], xk(k)); },
        [&](int k) { return cmul(weff_a(w, b, r, k, p, a2(k)), yk(k)); }));
  }
  return acc;
}

// ---------------------------------------------------------------- k_col
// 256 threads = 64 rows x 4 waves; wave s handles the panel terms k = s, s+4, ...
template <typename T>
__global__ void __launch_bounds__(256) k_col(TridiagWork<T> w, int64_t j, int jj, int64_t p) {
  __shared__ double sh[8];
  __shared__ cplx<T> vj[64], wj[64];     // row j of V and W_eff for k < jj
  __shared__ cplx<T> a2s[64];            // alpha2 of the panel columns k < jj
  __shared__ cplx<T> red[4][64];
  const int64_t b = blockIdx.z;
  const int64_t n = w.n;
  double* scal = w.scal + b * w.strideScal;
  double* part = w.part + b * w.strideP;
  cplx<T>* A = w.A + b * w.strideA;
  const cplx<T>* V = w.V + b * w.strideV;
  const int lane = threadIdx.x % 64, sub = threadIdx.x / 64;
  const int64_t r = j + blockIdx.x * (int64_t)64 + lane;
  // loads that do not wait for alpha2: this row's panel entries and A(r, j)
  PanelRow<T> prow;
  load_panel_row(w, b, r, sub, jj, p, r < n && sub < jj, prow);
  const cplx<T> a_rj = (sub == 0 && r < n) ? A[r + j * w.lda] : cplx<T>{0, 0};

  if (jj > 0) {
    // alpha2 of column j - 1, whose k_fin ran over m = n - j rows
    double a2r, a2i;
    alpha2_from_partials(part, w.yvOff, n - j, scal, sh, a2r, a2i);
    if (threadIdx.x == 0) {
      a2s[jj - 1] = cplx<T>{(T)a2r, (T)a2i};
      if (blockIdx.x == 0) { scal[S_A2 + 2 * (jj - 1)] = a2r; scal[S_A2 + 2 * (jj - 1) + 1] = a2i; }
    }
  }
  if ((int)threadIdx.x < jj - 1) {
    const int k = threadIdx.x;
    a2s[k] = cplx<T>{(T)scal[S_A2 + 2 * k], (T)scal[S_A2 + 2 * k + 1]};
  }
  __syncthreads();
  if ((int)threadIdx.x < jj) {
    const int k = threadIdx.x;
    vj[k] = V[j + (p + k) * n];
    wj[k] = weff_a(w, b, j, k, p, a2s[k]);
  }
  __syncthreads();

  cplx<T> acc{0, 0};
  if (r < n && sub < jj) {
    // V(r,k) conj(W(j,k)) + W(r,k) conj(V(j,k))
    acc = panel_row_terms<T>(
        w, b, r, sub, jj, p, prow, [&](int k) { return a2s[k]; },
        [&](int k) { return cconj(wj[k]); }, [&](int k) { return cconj(vj[k]); });
  }
  red[sub][lane] = acc;
  __syncthreads();
  double xpart = 0.0;
  if (sub == 0 && r < n) {
    cplx<T> a = a_rj;
    a = csub(a, cadd(cadd(red[0][lane], red[1][lane]), cadd(red[2][lane], red[3][lane])));
    if (r == j) a.y = 0;
    A[r + j * w.lda] = a;
    if (r >= j + 2) xpart = (double)cabs2(a);
  }
  xpart = block_sum(xpart, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = xpart;   // read by k_hemv_sym
}

// ---------------------------------------------------------------- k_hemv_sym
// y = S v for the Hermitian trailing matrix S = A(j+1:, j+1:), reading only
// its lower triangle once: each element of a 64 x 64 tile S_IJ (J <= I)
// contributes S_IJ v_J to y_I and S_IJ^H v_I to y_J.  grid: x = tile row I,
// y = segment of `seg` tiles along the row, z = batch.  Lane = row of the tile
// row, wave w = columns 16w..16w+15 of every tile; the next tile's elements
// are loaded while the current one is used.  Row sums stay in registers; the
// 16 column sums of a wave are reduced across lanes by a butterfly that halves
// the number of live columns at each step.  Results are stored as partials:
// the row sums of the workgroup in slot ys = blockIdx.y, the column sums of
// tile (I, J) in slot ntmax + I (each (slot, row) is written by exactly one
// workgroup).  Blocks of segment 0 also store the correction dots W^H x,
// V^H x of their tile row.
// Tiles start `sh` rows (and columns) before j + 1, at a multiple of 16: with
// lda a multiple of 16 every 64-row column segment then covers exactly four
// 128-byte lines instead of straddling five; local index l = o + sh for
// S-relative index o, valid for sh <= l < m + sh.
// most tiles per workgroup along a tile row (FP32: 4, FP64: 12, by measurement)
template <typename T> constexpr int kHemvSeg = sizeof(T) == 4 ? 4 : 12;

template <typename T>
__device__ __forceinline__ T xsel(bool h, T a, T b) { return h ? a : b; }

// Cross-lane steps of the butterfly without the LDS crossbar.
// v_permlane32_swap exchanges lanes 32..63 of x with lanes 0..31 of y, so
// afterwards x + y holds x(l) + x(l + 32) in lanes 0..31 and y(l - 32) + y(l)
// in lanes 32..63; v_permlane16_swap does the same for the odd 16-lane rows of
// x and the even rows of y.  64-bit values go through as two 32-bit halves.
template <int W32>
__device__ __forceinline__ void lane_swap(uint32_t& x, uint32_t& y) {
  auto r = W32 ? __builtin_amdgcn_permlane32_swap(x, y, false, false)
               : __builtin_amdgcn_permlane16_swap(x, y, false, false);
  x = r[0];
  y = r[1];
}
template <int W32>
__device__ __forceinline__ float fold_swap(float x, float y) {
  uint32_t a = __float_as_uint(x), b = __float_as_uint(y);
  lane_swap<W32>(a, b);
  return __uint_as_float(a) + __uint_as_float(b);
}
template <int W32>
__device__ __forceinline__ double fold_swap(double x, double y) {
  uint64_t a = __double_as_longlong(x), b = __double_as_longlong(y);
  uint32_t alo = (uint32_t)a, ahi = (uint32_t)(a >> 32), blo = (uint32_t)b, bhi = (uint32_t)(b >> 32);
  lane_swap<W32>(alo, blo);
  lane_swap<W32>(ahi, bhi);
  return __longlong_as_double((long long)(((uint64_t)ahi << 32) | alo)) +
         __longlong_as_double((long long)(((uint64_t)bhi << 32) | blo));
}
// DPP move within 16-lane rows: 0x128 row_ror:8 (partner l ^ 8), 0x141
// row_half_mirror (partner 7 - l within 8 lanes), 0x4E / 0xB1 quad
// permutations (partners l ^ 2, l ^ 1)
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __uint_as_float(__builtin_amdgcn_update_dpp(0u, __float_as_uint(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const uint64_t a = __double_as_longlong(v);
  const uint32_t lo = __builtin_amdgcn_update_dpp(0u, (uint32_t)a, CTRL, 0xf, 0xf, false);
  const uint32_t hi = __builtin_amdgcn_update_dpp(0u, (uint32_t)(a >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Butterfly reduce-scatter of 16 per-lane complex values over the 64 lanes:
// each step halves the live values, a lane keeping the half its bit selects.
// Afterwards the lanes with (lane & 3) == 0 hold the sum of value
// 8 h5 + 4 h4 + 2 h3 + h2 (hX = bit X of the lane index).
template <typename T>
__device__ __forceinline__ void reduce_scatter16(const T (&pr)[16], const T (&pi)[16], bool h3,
                                                 bool h2, T& r1, T& i1) {
  T r8[8], i8[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    r8[k] = fold_swap<1>(pr[k], pr[8 + k]);
    i8[k] = fold_swap<1>(pi[k], pi[8 + k]);
  }
  T r4[4], i4[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    r4[k] = fold_swap<0>(r8[k], r8[4 + k]);
    i4[k] = fold_swap<0>(i8[k], i8[4 + k]);
  }
  T r2[2], i2[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    r2[k] = xsel(h3, r4[2 + k], r4[k]) + dpp<0x128>(xsel(h3, r4[k], r4[2 + k]));
    i2[k] = xsel(h3, i4[2 + k], i4[k]) + dpp<0x128>(xsel(h3, i4[k], i4[2 + k]));
  }
  // the partner 7 - l has the other bit 2, so each quad of lanes still
  // covers all eight lanes of its group once the quad sums are taken
  r1 = xsel(h2, r2[1], r2[0]) + dpp<0x141>(xsel(h2, r2[0], r2[1]));
  i1 = xsel(h2, i2[1], i2[0]) + dpp<0x141>(xsel(h2, i2[0], i2[1]));
  r1 += dpp<0x4E>(r1);
  i1 += dpp<0x4E>(i1);
  r1 += dpp<0xB1>(r1);
  i1 += dpp<0xB1>(i1);
}

template <typename T>
__global__ void __launch_bounds__(256) k_hemv_sym(TridiagWork<T> w, int64_t j, int jj, int64_t p,
                                                  int seg, int sh) {
  __shared__ cplx<T> vseg[kHemvSeg<T> * 64];
  __shared__ cplx<T> red[4][64];
  // odd columns walk the batch backwards: the first workgroups then read the
  // matrices the previous launch read last, which the memory-side cache
  // (256 MB) may still hold
  const int64_t b = (j & 1) ? (int64_t)gridDim.z - 1 - blockIdx.z : (int64_t)blockIdx.z;
  const int64_t n = w.n, m = n - j - 1, mL = m + sh;
  const int64_t r0 = j + 1 - sh;                      // absolute row / column of local 0
  const int I = blockIdx.x;
  const int J0 = blockIdx.y * seg;
  if (J0 > I) return;
  const int J1 = min(I + 1, J0 + seg);
  double* scal = w.scal + b * w.strideScal;
  const cplx<T>* A = w.A + b * w.strideA;
  const cplx<T>* S = A + r0 + r0 * w.lda;
  const cplx<T>* V = w.V + b * w.strideV;
  cplx<T>* yrow = w.ypart + b * w.strideY + (int64_t)blockIdx.y * w.ldY;
  cplx<T>* ycol = w.ypart + b * w.strideY + (w.ntmax + I) * w.ldY;
  const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
  const int64_t rI = (int64_t)I * 64 + lane;          // this lane's local row
  const bool rowok = rI >= sh && rI < mL;

  cplx<T> cur[16], nxt[16];
  auto load_tile = [&](int J, cplx<T>(&dst)[16]) {
#pragma unroll
    for (int cc = 0; cc < 16; ++cc) {
      const int64_t col = (int64_t)J * 64 + wv * 16 + cc;
      const bool ok = rowok && col >= sh && col < mL;
      const cplx<T> a = S[(ok ? rI : sh) + (ok ? col : sh) * w.lda];
      dst[cc] = ok ? a : cplx<T>{0, 0};
    }
  };
  load_tile(J0, nxt);
  // x = (0, A(j+2:, j)): the reflector's direction before scaling (local index)
  auto xval = [&](int64_t l) -> cplx<T> {
    return (l <= sh || l >= mL) ? cplx<T>{0, 0} : A[(r0 + l) + j * w.lda];
  };
  const cplx<T> vI = xval(rI);
  for (int o = tid; o < (J1 - J0) * 64; o += blockDim.x) vseg[o] = xval((int64_t)J0 * 64 + o);
  if (blockIdx.x == 0 && blockIdx.y == 0) {
    // the reflector of column j, from k_col's partial norms (k_col ran over n - j rows)
    __shared__ double sh[8];
    const double xn2 = sum_partials(w.part + b * w.strideP, (n - j + 63) / 64, 1, sh);
    if (tid == 0) {
      const cplx<T> ajj = A[j + j * w.lda], al = A[(j + 1) + j * w.lda];
      const Reflector rf = make_reflector(xn2, al.x, al.y);
      w.d[b * w.strideDE + j] = ajj.x;
      w.e[b * w.strideDE + j] = (T)rf.beta;
      w.tau[b * w.strideTau + j] = cplx<T>{(T)rf.tr, (T)rf.ti};
      scal[S_TAU] = rf.tr; scal[S_TAU + 1] = rf.ti;
      scal[S_SCALE] = rf.sr; scal[S_SCALE + 1] = rf.si;
    }
  }
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
    // the 16 column sums of the wave, reduced across its lanes
    T r1, i1;
    reduce_scatter16(pr, pi, h3, h2, r1, i1);
    const int64_t cJ = (int64_t)J * 64 + mycol;
    if ((lane & 3) == 0 && cJ >= sh && cJ < mL) ycol[cJ] = cplx<T>{r1, i1};
  }
  red[wv][lane] = rowacc;
  __syncthreads();
  if (wv == 0 && rowok)
    yrow[rI] = cadd(cadd(red[0][lane], red[1][lane]), cadd(red[2][lane], red[3][lane]));
  if (blockIdx.y == 0 && jj > 0) {
    // correction dots W_eff^H x (k < jj), V^H x (jj <= k < 2 jj) over this
    // tile row, 64 at a time (2 jj < 128): wave wv takes k = k0 + wv + 4 i,
    // i < 16; every load of a chunk is issued before its products are reduced
    // together (these workgroups run longest, and the last batch's ones end
    // the launch)
    const cplx<T> vr = rowok ? vI : cplx<T>{0, 0};
    const int64_t rr = r0 + (rowok ? rI : (int64_t)sh);     // a valid row when masked
    for (int k0 = 0; k0 < 2 * jj; k0 += 64) {
      T dr[16], di[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int k = k0 + wv + 4 * i;
        const bool ok = k < 2 * jj;
        const int kc = ok ? k : 0;
        const cplx<T> x = (kc < jj) ? weff(w, b, rr, kc, p, scal) : V[rr + (p + kc - jj) * n];
        const cplx<T> t = cmulc(x, vr);
        dr[i] = ok ? t.x : T(0);
        di[i] = ok ? t.y : T(0);
      }
      T r1, i1;
      reduce_scatter16(dr, di, h3, h2, r1, i1);
      const int kk = k0 + wv + 4 * ((h5 ? 8 : 0) + (h4 ? 4 : 0) + (h3 ? 2 : 0) + (h2 ? 1 : 0));
      if ((lane & 3) == 0 && kk < 2 * jj)
        w.dpart[b * w.strideDp + (int64_t)I * 2 * w.nb + kk] = cplx<T>{r1, i1};
    }
  }
}

// ---------------------------------------------------------------- k_fin
// 256 threads = 64 rows x 4 waves; wave s handles the panel terms k = s, s+4, ...
template <typename T>
__global__ void __launch_bounds__(256) k_fin(TridiagWork<T> w, int64_t j, int jj, int64_t p,
                                             int seg, int tsh) {
  __shared__ cplx<T> sW[64], sV[64];
  __shared__ cplx<T> red[4][64];
  __shared__ double ured[4][64][2], dred[4][128][2];
  __shared__ double sh[8];
  const int64_t b = blockIdx.z;
  const int64_t n = w.n, m = n - j - 1;
  double* scal = w.scal + b * w.strideScal;
  cplx<T>* V = w.V + b * w.strideV;
  const cplx<T>* A = w.A + b * w.strideA;
  const cplx<T> tau{(T)scal[S_TAU], (T)scal[S_TAU + 1]};
  const cplx<T> sc{(T)scal[S_SCALE], (T)scal[S_SCALE + 1]};
  const int lane = threadIdx.x % 64, sub = threadIdx.x / 64;
  const int64_t rr = blockIdx.x * (int64_t)64 + lane;
  const int64_t r = j + 1 + rr;
  // loads that do not wait for the dots: this row's panel entries, S(r, 0),
  // u(r) and x(r) = A(r, j)
  PanelRow<T> prow;
  load_panel_row(w, b, r, sub, jj, p, rr < m && sub < jj, prow);
  cplx<T> s0{0, 0}, xr{0, 0};
  if (sub == 0 && rr < m) {
    s0 = A[r + (j + 1) * w.lda];
    xr = A[r + j * w.lda];
  }
  // u(r) from the hemv's partials of local row l = rr + tsh in tile row tr: the
  // row sums of segments 0 .. tr / seg and the column sums of tile rows
  // tr .. nt - 1; wave `sub` takes terms sub, sub + 4, ... with eight loads in
  // flight, summed in FP64 in a fixed order
  const int64_t nt = (m + tsh + 63) / 64;
  {
    const int64_t l = rr + tsh, tr = l / 64;
    const int64_t nrow = tr / seg + 1, nterm = (rr < m) ? nrow + (nt - tr) : 0;
    const cplx<T>* P = w.ypart + b * w.strideY + (rr < m ? l : 0);
    double ux = 0.0, uy = 0.0;
    for (int64_t t0 = sub; t0 < nterm; t0 += 32) {
      cplx<T> v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int64_t t = min(t0 + 4 * u, nterm - 1);
        const int64_t slot = (t < nrow) ? t : w.ntmax + tr + (t - nrow);
        v[u] = P[slot * w.ldY];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (t0 + 4 * u < nterm) { ux += (double)v[u].x; uy += (double)v[u].y; }
    }
    ured[sub][lane][0] = ux;
    ured[sub][lane][1] = uy;
  }
  // the correction dots W^H x, V^H x (2 jj <= 128 of them) summed over the
  // tile rows: dot k = lane, lane + 64; wave `sub` takes tile rows sub,
  // sub + 4, ..., the same way
  for (int k = lane; k < 2 * jj; k += 64) {
    const cplx<T>* D = w.dpart + b * w.strideDp + k;
    double dx = 0.0, dy = 0.0;
    for (int64_t I0 = sub; I0 < nt; I0 += 32) {
      cplx<T> v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = D[min(I0 + 4 * u, nt - 1) * 2 * w.nb];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (I0 + 4 * u < nt) { dx += (double)v[u].x; dy += (double)v[u].y; }
    }
    dred[sub][k][0] = dx;
    dred[sub][k][1] = dy;
  }
  __syncthreads();
  // W^H v = conj(W_eff(j+1, k)) + scale (W^H x), V^H v likewise
  if (threadIdx.x < 2 * jj) {
    const int k = threadIdx.x;
    const cplx<T> sdot{(T)(((dred[0][k][0] + dred[1][k][0]) + dred[2][k][0]) + dred[3][k][0]),
                       (T)(((dred[0][k][1] + dred[1][k][1]) + dred[2][k][1]) + dred[3][k][1])};
    if (k < jj) sW[k] = cadd(cconj(weff(w, b, j + 1, k, p, scal)), cmul(sc, sdot));
    else sV[k - jj] = cadd(cconj(V[(j + 1) + (p + k - jj) * n]), cmul(sc, sdot));
  }
  __syncthreads();
  const cplx<T> ur{(T)(((ured[0][lane][0] + ured[1][lane][0]) + ured[2][lane][0]) + ured[3][lane][0]),
                   (T)(((ured[0][lane][1] + ured[1][lane][1]) + ured[2][lane][1]) + ured[3][lane][1])};

  cplx<T> acc{0, 0};
  if (rr < m && sub < jj) {
    acc = panel_row_terms<T>(
        w, b, r, sub, jj, p, prow,
        [&](int k) { return cplx<T>{(T)scal[S_A2 + 2 * k], (T)scal[S_A2 + 2 * k + 1]}; },
        [&](int k) { return sW[k]; }, [&](int k) { return sV[k]; });
  }
  red[sub][lane] = acc;
  __syncthreads();
  double pr = 0, pi = 0;
  if (sub == 0 && rr < m) {
    // y = S(:, 0) + scale u, S(:, 0) = A(j+1:, j+1) with a real diagonal
    if (rr == 0) s0.y = 0;
    cplx<T> y = cadd(s0, cmul(sc, ur));
    y = csub(y, cadd(cadd(red[0][lane], red[1][lane]), cadd(red[2][lane], red[3][lane])));
    y = cmul(tau, y);
    w.W[b * w.strideW + r + (int64_t)jj * n] = y;
    const cplx<T> v = (rr == 0) ? cplx<T>{1, 0} : cmul(sc, xr);
    V[r + j * n] = v;
    const cplx<T> t = cmulc(y, v);
    pr = t.x; pi = t.y;
  }
  // per-workgroup partials of w^H v; alpha2 = -tau/2 (w^H v) is formed by
  // the next k_col (or k_fixw at the end of the panel)
  pr = block_sum(pr, sh);
  pi = block_sum(pi, sh);
  if (threadIdx.x == 0) {
    double* part = w.part + b * w.strideP + w.yvOff;
    part[2 * blockIdx.x] = pr;
    part[2 * blockIdx.x + 1] = pi;
  }
}

// ---------------------------------------------------------------- panel end
// W(r, k) <- W_eff(r, k) for rows >= r0 and k < kw; W(:, k) = 0 for kw <= k < kpad.
template <typename T>
__global__ void __launch_bounds__(256) k_fixw(TridiagWork<T> w, int64_t r0, int kw, int kpad,
                                              int64_t p) {
  __shared__ double sh[8];
  const int64_t b = blockIdx.z;
  const double* scal = w.scal + b * w.strideScal;
  const int64_t r = r0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  cplx<T> a2{0, 0};
  if (k == kw - 1) {
    // alpha2 of the panel's last column, whose k_fin ran over n - r0 rows
    double a2r, a2i;
    alpha2_from_partials(w.part + b * w.strideP, w.yvOff, w.n - r0, scal, sh, a2r, a2i);
    a2 = cplx<T>{(T)a2r, (T)a2i};
  } else if (k < kw) {
    a2 = cplx<T>{(T)scal[S_A2 + 2 * k], (T)scal[S_A2 + 2 * k + 1]};
  }
  if (r >= w.n || k >= kpad) return;
  cplx<T>* Wc = w.W + b * w.strideW + r + (int64_t)k * w.n;
  if (k < kw) *Wc = weff_a(w, b, r, k, p, a2);
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
}

template <typename T>
void tridiagonalise(const TridiagWork<T>& w, hipStream_t s) {
  const int64_t n = w.n;
  const unsigned B = (unsigned)w.batch;
  if (w.nb > 64 || w.nb % 16 != 0) throw std::runtime_error("tridiagonalise: nb must be 16..64, multiple of 16");
  if (n > 1 && (w.ntmax < hemv_tiles(n) || w.ldY < w.ntmax * 64 || w.strideY < 2 * w.ntmax * w.ldY ||
                w.strideDp < w.ntmax * 2 * w.nb))
    throw std::runtime_error("tridiagonalise: hemv partial buffers too small");
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
      // tiles start at a multiple of 16 rows when lda allows 128-byte alignment
      const int sh = (w.lda % 16 == 0) ? (int)((j + 1) % 16) : 0;
      const int64_t nt = ceil_div(m + sh, 64);
      // tiles per workgroup: up to kHemvSeg, fewer once the batch has less
      // than 2048 kHemvSeg lower-triangle tiles, so that small trailing
      // matrices are not left to a few serial chains (measured on 16 x 4764^2
      // FP32: 4 tiles stream 5.2 TB/s at m = 4763 where 12 give 5.0; 1-2 are
      // best below m ~ 1200)
      const int64_t tiles = nt * (nt + 1) / 2 * (int64_t)B;
      const int seg = (int)std::max<int64_t>(1, std::min<int64_t>(kHemvSeg<T>, tiles / 2048));
      const bool timed = w.hemv_ev && (j % kHemvTimeStride == 0);
      if (timed) PPH_HIP(hipEventRecord(w.hemv_ev[2 * (j / kHemvTimeStride)], s));
      hipLaunchKernelGGL(k_hemv_sym<T>, dim3((unsigned)nt, (unsigned)ceil_div(nt, seg), B), dim3(256),
                         0, s, w, j, jj, p, seg, sh);
      PPH_CHECK_LAUNCH();
      if (timed) PPH_HIP(hipEventRecord(w.hemv_ev[2 * (j / kHemvTimeStride) + 1], s));
      hipLaunchKernelGGL(k_fin<T>, dim3((unsigned)ceil_div(m, 64), 1, B), dim3(256), 0, s,
                         w, j, jj, p, seg, sh);
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
