// Complex GEMM on CDNA4 matrix cores (MFMA), FP64 and FP32, with a fused
// "diagonal dot" epilogue used by the group-velocity and Rayleigh-quotient
// products (sum_rows Re(conj(X) .* (A B))) so the product is never stored.
//
// A complex product takes three real MFMAs per 16x16x4 step (the 3M scheme):
// P1 += Ar Br, P2 += Ai Bi, P3 += (Ar + Ai)(Br + Bi), Re = P1 - P2,
// Im = P3 - P1 - P2.  The kernel is held to two waves per SIMD so that the FP64
// instance keeps its three accumulator sets in VGPRs without losing occupancy
// (the four-MFMA form, P1 = Re and P2 = Im, stays selectable through k3M).
// Tile: 64x64 per 256-thread workgroup (2x2 waves, each 32x32 = 2x2 MFMA
// blocks), K staged through LDS in steps of 16.
#include "common.hpp"
#include "kernels.hpp"

namespace pph {

constexpr int GBK = 16;

// WM x WN MFMA blocks of 16 x 16 per wave: the workgroup tile is
// (32 WM) x (32 WN); the library dispatches 2 x 2 (64 x 64)
template <typename T, bool CONJA, bool CONJB, bool DOT, int WM = 2, int WN = 2, bool M3 = true>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WM * WN > 4 ? 1 : 2)))
k_zgemm(GemmParams<T> p) {
  using acc_t = typename mfma<T>::acc;
  constexpr int GBM = 32 * WM, GBN = 32 * WN;
  // two LDS stages (the next K tile is stored while the current one is
  // consumed: one barrier per K step); rows padded by 16 bytes so the
  // k-major stores of 16 lanes fall in distinct banks
  constexpr int PAD = 16 / (int)sizeof(cplx<T>);
  __shared__ cplx<T> As[2][GBK][GBM + PAD];
  __shared__ cplx<T> Bs[2][GBK][GBN + PAD];

  const int tid = threadIdx.x;
  const int lane = tid % WAVE, wave = tid / WAVE;
  const int wr0 = (wave / 2) * 16 * WM, wc0 = (wave % 2) * 16 * WN;
  const int64_t ksplit = p.ksplit > 1 ? p.ksplit : 1;
  const int64_t binner = p.binner > 1 ? p.binner : 1;
  int64_t z = blockIdx.z;
  const int64_t ks_idx = z % ksplit;
  z /= ksplit;
  const int64_t bi = z % binner, b = z / binner;
  const int64_t row0 = (int64_t)blockIdx.x * GBM, col0 = (int64_t)blockIdx.y * GBN;
  if (p.lower && col0 > row0 + GBM - 1) return;
  const int64_t offA = b * p.strideA + bi * p.sAin, offB = b * p.strideB + bi * p.sBin;
  const cplx<T>* A1 = p.A + offA;
  const cplx<T>* B1 = p.B + offB;
  const bool two = (p.A2 != nullptr);
  const cplx<T>* A2 = two ? p.A2 + b * p.strideA2 : A1;
  const cplx<T>* B2 = two ? p.B2 + b * p.strideB2 : B1;
  const int64_t M = p.m, N = p.n;
  // K range of this split
  int64_t kbeg = 0, K = p.k;
  if (ksplit > 1) {
    const int64_t kc = ceil_div_dev(ceil_div_dev(p.k, ksplit), GBK) * GBK;
    kbeg = ks_idx * kc;
    K = min(p.k, kbeg + kc);
  }
  // Hermitian dot: A'(r, k) = 0 for k > r, so the row tile needs k < row0 + GBM;
  // triangular op(A) likewise limits the K range of the row tile
  const bool herm = p.herm_dot != 0;
  if (herm || p.tri == 1) K = min(K, row0 + GBM);
  if (p.tri == 2) kbeg = max(kbeg, (row0 / GBK) * GBK);
  const int64_t K1 = two ? p.k1 : K;
  const bool tri1 = p.tri == 1, tri2 = p.tri == 2;

  constexpr bool k3M = M3;
  auto product = [](T p1, T p2, T p3) {
    return k3M ? cplx<T>{p1 - p2, p3 - p1 - p2} : cplx<T>{p1, p2};
  };
  acc_t acc1[WM][WN], acc2[WM][WN], acc3[WM][WN];
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      acc1[i][j] = acc_t{0, 0, 0, 0};
      acc2[i][j] = acc_t{0, 0, 0, 0};
      acc3[i][j] = acc_t{0, 0, 0, 0};
    }

  cplx<T> ra[2 * WM], rb[2 * WN];
  // Operand tiles are loaded unconditionally from clamped addresses into
  // registers; masking, conjugation and the Hermitian-dot diagonal rule are
  // applied only when the registers are stored to LDS, after the current K
  // step's MFMAs.  No loaded value is used before then, so the next tile's
  // loads stay in flight across the whole K step instead of being waited for
  // at its start.
  int64_t kpend = kbeg;                 // K offset of the tile held in ra / rb
  struct Seg { const cplx<T>* A; const cplx<T>* B; int64_t k0, Kseg; };
  auto seg = [&](int64_t kg) {
    const bool second = two && (kg >= K1);
    return Seg{second ? A2 : A1, second ? B2 : B1, second ? kg - K1 : kg, second ? K - K1 : K1};
  };
  auto load_tiles = [&](int64_t kg) {
    kpend = kg;
    const Seg g = seg(kg);
#pragma unroll
    for (int i = 0; i < 2 * WM; ++i) {
      if (!CONJA) {
        const int r = tid % GBM, kk = tid / GBM + (256 / GBM) * i;
        const int64_t gr = row0 + r, gk = g.k0 + kk;
        const bool ok = gr < M && gk < g.Kseg;
        ra[i] = g.A[ok ? gr + gk * p.lda : 0];
      } else {
        const int kk = tid % 16, r = tid / 16 + 16 * i;
        const int64_t gr = row0 + r, gk = g.k0 + kk;
        const bool ok = gr < M && gk < g.Kseg;
        ra[i] = g.A[ok ? gk + gr * p.lda : 0];
      }
    }
#pragma unroll
    for (int i = 0; i < 2 * WN; ++i) {
      if (!CONJB) {
        const int kk = tid % 16, c = tid / 16 + 16 * i;
        const int64_t gk = g.k0 + kk, gc = col0 + c;
        const bool ok = gk < g.Kseg && gc < N;
        rb[i] = g.B[ok ? gk + gc * p.ldb : 0];
      } else {
        const int c = tid % GBN, kk = tid / GBN + (256 / GBN) * i;
        const int64_t gk = g.k0 + kk, gc = col0 + c;
        const bool ok = gk < g.Kseg && gc < N;
        rb[i] = g.B[ok ? gc + gk * p.ldb : 0];
      }
    }
  };
  auto mask = [](cplx<T> v, bool ok, bool conj) {
    const T zr = T(0);
    return cplx<T>{ok ? v.x : zr, ok ? (conj ? -v.y : v.y) : zr};
  };
  auto store_tiles = [&](int st) {
    const Seg g = seg(kpend);
#pragma unroll
    for (int i = 0; i < 2 * WM; ++i) {
      if (!CONJA) {
        const int r = tid % GBM, kk = tid / GBM + (256 / GBM) * i;
        const int64_t gr = row0 + r, gk = g.k0 + kk;
        cplx<T> v = mask(ra[i], gr < M && gk < g.Kseg && !(tri1 && gk > gr) &&
                                    !(tri2 && gk <= gr), false);
        if (herm && gk >= gr) v = (gk == gr) ? cplx<T>{T(0.5) * v.x, T(0)} : cplx<T>{0, 0};
        As[st][kk][r] = v;
      } else {
        const int kk = tid % 16, r = tid / 16 + 16 * i;
        const int64_t gr = row0 + r, gk = g.k0 + kk;
        As[st][kk][r] = mask(ra[i], gr < M && gk < g.Kseg && !(tri1 && gk > gr) &&
                                        !(tri2 && gk <= gr), true);
      }
    }
#pragma unroll
    for (int i = 0; i < 2 * WN; ++i) {
      if (!CONJB) {
        const int kk = tid % 16, c = tid / 16 + 16 * i;
        const int64_t gk = g.k0 + kk, gc = col0 + c;
        Bs[st][kk][c] = mask(rb[i], gk < g.Kseg && gc < N, false);
      } else {
        const int c = tid % GBN, kk = tid / GBN + (256 / GBN) * i;
        const int64_t gk = g.k0 + kk, gc = col0 + c;
        Bs[st][kk][c] = mask(rb[i], gk < g.Kseg && gc < N, true);
      }
    }
  };

  load_tiles(kbeg);
  store_tiles(0);
  __syncthreads();
  int st = 0;
  for (int64_t k0 = kbeg; k0 < K; k0 += GBK) {
    const bool more = k0 + GBK < K;
    if (more) load_tiles(k0 + GBK);
#pragma unroll
    for (int ks = 0; ks < GBK / 4; ++ks) {
      const int kk = ks * 4 + lane / 16;
      cplx<T> a[WM], bb[WN];
#pragma unroll
      for (int mi = 0; mi < WM; ++mi) a[mi] = As[st][kk][wr0 + mi * 16 + lane % 16];
#pragma unroll
      for (int ni = 0; ni < WN; ++ni) bb[ni] = Bs[st][kk][wc0 + ni * 16 + lane % 16];
      T as_[WM] = {}, bs_[WN] = {};
      if constexpr (k3M) {
#pragma unroll
        for (int mi = 0; mi < WM; ++mi) as_[mi] = a[mi].x + a[mi].y;
#pragma unroll
        for (int ni = 0; ni < WN; ++ni) bs_[ni] = bb[ni].x + bb[ni].y;
      }
#pragma unroll
      for (int mi = 0; mi < WM; ++mi)
#pragma unroll
        for (int ni = 0; ni < WN; ++ni) {
          if constexpr (k3M) {
            acc1[mi][ni] = mfma<T>::op(a[mi].x, bb[ni].x, acc1[mi][ni]);
            acc2[mi][ni] = mfma<T>::op(a[mi].y, bb[ni].y, acc2[mi][ni]);
            acc3[mi][ni] = mfma<T>::op(as_[mi], bs_[ni], acc3[mi][ni]);
          } else {
            acc1[mi][ni] = mfma<T>::op(a[mi].x, bb[ni].x, acc1[mi][ni]);
            acc1[mi][ni] = mfma<T>::op(-a[mi].y, bb[ni].y, acc1[mi][ni]);
            acc2[mi][ni] = mfma<T>::op(a[mi].x, bb[ni].y, acc2[mi][ni]);
            acc2[mi][ni] = mfma<T>::op(a[mi].y, bb[ni].x, acc2[mi][ni]);
          }
        }
    }
    if (more) {
      // the other stage was last read before the previous barrier
      store_tiles(st ^ 1);
      st ^= 1;
    }
    __syncthreads();
  }

  if (!DOT) {
    // a split-K piece writes its own slice (launch_gemm points C at the workspace)
    cplx<T>* C = p.C + b * p.strideC + bi * p.sCin + ks_idx * p.sCsplit;
    const bool beta0 = (p.beta.x == T(0) && p.beta.y == T(0));
#pragma unroll
    for (int mi = 0; mi < WM; ++mi)
#pragma unroll
      for (int ni = 0; ni < WN; ++ni) {
        const int64_t col = col0 + wc0 + ni * 16 + lane % 16;
        if (col >= N) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int64_t row = row0 + wr0 + mi * 16 + mfma<T>::row(lane, i);
          if (row >= M) continue;
          cplx<T> v = product(acc1[mi][ni][i], acc2[mi][ni][i], acc3[mi][ni][i]);
          v = cmul(p.alpha, v);
          cplx<T>* c = C + row + col * p.ldc;
          if (!beta0) v = cadd(v, cmul(p.beta, *c));
          *c = v;
        }
      }
  } else {
    // per column: the two waves of a column half add their rows in LDS (upper
    // half first), and the row tile's sum goes to its partial slot
    __shared__ double dlow[GBN];
    const cplx<T>* X = p.dotX + b * p.strideX + bi * p.sXin;
    double* dpart = static_cast<double*>(p.ws) + (z * (int64_t)gridDim.x + blockIdx.x) * N;
    double dsum[WN];
#pragma unroll
    for (int ni = 0; ni < WN; ++ni) {
      const int64_t col = col0 + wc0 + ni * 16 + lane % 16;
      double s = 0.0;
      if (col < N) {
#pragma unroll
        for (int mi = 0; mi < WM; ++mi)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int64_t row = row0 + wr0 + mi * 16 + mfma<T>::row(lane, i);
            if (row >= M) continue;
            const cplx<T> x = X[row + col * p.lddx];
            const cplx<T> v = product(acc1[mi][ni][i], acc2[mi][ni][i], acc3[mi][ni][i]);
            // Re(conj(x) * acc)
            s += (double)x.x * (double)v.x + (double)x.y * (double)v.y;
          }
      }
      if (herm) s *= 2.0;
      s += __shfl_xor(s, 16, WAVE);
      s += __shfl_xor(s, 32, WAVE);
      dsum[ni] = s;
      if (wave >= 2 && lane < 16) dlow[wc0 + ni * 16 + lane] = s;
    }
    __syncthreads();
    if (wave < 2 && lane < 16) {
#pragma unroll
      for (int ni = 0; ni < WN; ++ni) {
        const int c = wc0 + ni * 16 + lane;
        if (col0 + c < N) dpart[col0 + c] = dsum[ni] + dlow[c];
      }
    }
  }
}

// out[(b binner + bi) ...] += sum over the row tiles of the dot partials, in order
__global__ void __launch_bounds__(256) k_dot_reduce(const double* __restrict__ part, int64_t rt,
                                                    int64_t N, int64_t binner, double* dout,
                                                    int64_t dot_stride, int64_t sDotin) {
  const int64_t col = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t z = blockIdx.y;
  if (col >= N) return;
  const double* P = part + z * rt * N + col;
  double s = 0.0;
  for (int64_t t = 0; t < rt; ++t) s += P[t * N];
  dout[(z / binner) * dot_stride + (z % binner) * sDotin + col] += s;
}

// C = sum over the split-K pieces (in order) + beta C
template <typename T>
__global__ void __launch_bounds__(256) k_split_reduce(const cplx<T>* __restrict__ W, int64_t ks,
                                                      int64_t M, int64_t N, int64_t binner,
                                                      cplx<T>* C, int64_t ldc, int64_t strideC,
                                                      int64_t sCin, cplx<T> beta) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t z = blockIdx.y;
  if (idx >= M * N) return;
  const int64_t row = idx % M, col = idx / M;
  const cplx<T>* P = W + z * ks * M * N + idx;
  cplx<T> v{0, 0};
  for (int64_t k = 0; k < ks; ++k) v = cadd(v, P[k * M * N]);
  cplx<T>* c = C + (z / binner) * strideC + (z % binner) * sCin + row + col * ldc;
  if (beta.x != T(0) || beta.y != T(0)) v = cadd(v, cmul(beta, *c));
  *c = v;
}

template <typename T>
int64_t gemm_ws_bytes(const GemmParams<T>& p) {
  const int64_t bin = p.binner > 1 ? p.binner : 1;
  if (p.m <= 0 || p.n <= 0 || p.batch <= 0) return 0;
  if (p.dot_out != nullptr) return ceil_div(p.m, 64) * p.batch * bin * p.n * (int64_t)sizeof(double);
  if (p.ksplit > 1) return p.ksplit * p.batch * bin * p.m * p.n * (int64_t)sizeof(cplx<T>);
  return 0;
}

template <typename T> void launch_gemm_tiles(const GemmParams<T>& p, hipStream_t s);

template <typename T>
void launch_gemm(const GemmParams<T>& p, hipStream_t s) {
  if (p.m <= 0 || p.n <= 0 || p.batch <= 0) return;
  const int64_t ks = p.ksplit > 1 ? p.ksplit : 1, bin = p.binner > 1 ? p.binner : 1;
  if (ks > 1 && p.A2 != nullptr) throw std::runtime_error("launch_gemm: split-K with two products");
  if (p.herm_dot && (p.dot_out == nullptr || p.opA != 'N' || p.A2 != nullptr || p.m != p.k))
    throw std::runtime_error("launch_gemm: herm_dot needs the dot epilogue, opA = 'N', square A, one product");
  if (p.tri && (p.dot_out != nullptr || p.A2 != nullptr || p.ksplit > 1))
    throw std::runtime_error("launch_gemm: a triangular A takes one product, no dot epilogue, no split-K");
  if (p.dot_out != nullptr && (ks > 1 || p.lower))
    throw std::runtime_error("launch_gemm: the dot epilogue takes no split-K and no lower-only tiles");
  if (gemm_ws_bytes(p) > 0 && p.ws == nullptr)
    throw std::runtime_error("launch_gemm: the dot epilogue and split-K need a workspace");
  if (ks > 1) {
    // the pieces into the workspace, [(b binner + bi) ksplit + piece] x (M x N, ld M)
    GemmParams<T> q = p;
    q.C = static_cast<cplx<T>*>(p.ws);
    q.ldc = p.m;
    q.sCsplit = p.m * p.n;
    q.sCin = ks * q.sCsplit;
    q.strideC = bin * q.sCin;
    q.beta = cplx<T>{0, 0};
    launch_gemm_tiles(q, s);
  } else {
    launch_gemm_tiles(p, s);
  }
  if (p.dot_out != nullptr) {
    dim3 g((unsigned)ceil_div(p.n, 256), (unsigned)(p.batch * bin));
    hipLaunchKernelGGL(k_dot_reduce, g, dim3(256), 0, s, static_cast<const double*>(p.ws),
                       ceil_div(p.m, 64), p.n, bin, p.dot_out, p.dot_stride, p.sDotin);
    PPH_CHECK_LAUNCH();
  } else if (ks > 1) {
    dim3 g((unsigned)ceil_div(p.m * p.n, 256), (unsigned)(p.batch * bin));
    hipLaunchKernelGGL(k_split_reduce<T>, g, dim3(256), 0, s, static_cast<const cplx<T>*>(p.ws), ks,
                       p.m, p.n, bin, p.C, p.ldc, p.strideC, p.sCin, p.beta);
    PPH_CHECK_LAUNCH();
  }
}

// ---------------------------------------------------------------- tile shapes
// The shape a product with tile 0 takes: the calling thread's scope, else the
// process default (set by pph_debug_gemm_tile with no handle), else 64 x 64.
static int g_gemm_tile_default = kGemmTile64x64;
static thread_local int t_gemm_tile = kGemmTileDefault;

void set_gemm_tile_default(int tile) { g_gemm_tile_default = tile; }
int gemm_current_tile() { return t_gemm_tile != kGemmTileDefault ? t_gemm_tile : g_gemm_tile_default; }
gemm_tile_scope::gemm_tile_scope(int tile) : saved(t_gemm_tile) { t_gemm_tile = tile; }
gemm_tile_scope::~gemm_tile_scope() { t_gemm_tile = saved; }

// (32 WM) x (32 WN) tiles, three or four real products per complex step
template <typename T, bool CA, bool CB, int WM, int WN>
static void launch_shape(const GemmParams<T>& p, int64_t zdim, hipStream_t s) {
  const dim3 g((unsigned)ceil_div(p.m, 32 * WM), (unsigned)ceil_div(p.n, 32 * WN), (unsigned)zdim);
  if (p.four_m)
    hipLaunchKernelGGL((k_zgemm<T, CA, CB, false, WM, WN, false>), g, dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL((k_zgemm<T, CA, CB, false, WM, WN>), g, dim3(256), 0, s, p);
  PPH_CHECK_LAUNCH();
}

template <typename T, bool CA, bool CB>
static void launch_tiled(const GemmParams<T>& p, int tile, int64_t zdim, hipStream_t s) {
  switch (tile) {
    case kGemmTile128x64: launch_shape<T, CA, CB, 4, 2>(p, zdim, s); break;
    case kGemmTile64x128: launch_shape<T, CA, CB, 2, 4>(p, zdim, s); break;
    case kGemmTile128x128:
      // FP64 accumulators of a 128 x 128 tile do not fit the register file
      if constexpr (sizeof(T) == 8) throw std::runtime_error("launch_gemm: the 128 x 128 tile is FP32 only");
      else launch_shape<T, CA, CB, 4, 4>(p, zdim, s);
      break;
    default: launch_shape<T, CA, CB, 2, 2>(p, zdim, s); break;
  }
}

// the tile kernels of launch_gemm (arguments checked by the caller)
template <typename T>
void launch_gemm_tiles(const GemmParams<T>& p, hipStream_t s) {
  const int64_t ks = p.ksplit > 1 ? p.ksplit : 1, bin = p.binner > 1 ? p.binner : 1;
  // a K step must not straddle the two products
  if (p.A2 != nullptr && p.k1 % GBK != 0) throw std::runtime_error("launch_gemm: k1 must be a multiple of 16");
  const bool ca = (p.opA == 'C'), cb = (p.opB == 'C'), dot = (p.dot_out != nullptr);
  // outputs at most 32 columns wide (the panel products of the band
  // reduction) take 64 x 32 tiles, so no MFMA block is spent on empty columns
  if (p.n <= 32 && !dot && !p.lower) {
    dim3 g32((unsigned)ceil_div(p.m, 64), 1u, (unsigned)(p.batch * bin * ks));
#define PPH_GEMM_CASE32(CA, CB)                                                    \
    if (ca == CA && cb == CB) {                                                    \
      if (p.four_m)                                                                \
        hipLaunchKernelGGL((k_zgemm<T, CA, CB, false, 2, 1, false>), g32, dim3(256), 0, s, p); \
      else                                                                         \
        hipLaunchKernelGGL((k_zgemm<T, CA, CB, false, 2, 1>), g32, dim3(256), 0, s, p); \
      PPH_CHECK_LAUNCH();                                                          \
      return;                                                                      \
    }
    PPH_GEMM_CASE32(false, false)
    PPH_GEMM_CASE32(true, false)
    PPH_GEMM_CASE32(false, true)
    PPH_GEMM_CASE32(true, true)
#undef PPH_GEMM_CASE32
  }
  const int64_t zdim = p.batch * bin * ks;
  if (!dot) {
    const int tile = p.tile != kGemmTileDefault ? p.tile : gemm_current_tile();
    if (tile < kGemmTile64x64 || tile >= kGemmTileCount)
      throw std::runtime_error("launch_gemm: unknown tile shape " + std::to_string(tile));
    if (!ca && !cb) return launch_tiled<T, false, false>(p, tile, zdim, s);
    if (ca && !cb) return launch_tiled<T, true, false>(p, tile, zdim, s);
    if (!ca && cb) return launch_tiled<T, false, true>(p, tile, zdim, s);
    return launch_tiled<T, true, true>(p, tile, zdim, s);
  }
  // the dot epilogue: 64 x 64 tiles (its partial slots are per 64-row tile)
  if (ca || cb) throw std::runtime_error("launch_gemm: the dot epilogue takes op 'N' for both operands");
  const dim3 grid((unsigned)ceil_div(p.m, 64), (unsigned)ceil_div(p.n, 64), (unsigned)zdim);
  hipLaunchKernelGGL((k_zgemm<T, false, false, true>), grid, dim3(256), 0, s, p);
  PPH_CHECK_LAUNCH();
}

template void launch_gemm<double>(const GemmParams<double>&, hipStream_t);
template void launch_gemm<float>(const GemmParams<float>&, hipStream_t);
template int64_t gemm_ws_bytes<double>(const GemmParams<double>&);
template int64_t gemm_ws_bytes<float>(const GemmParams<float>&);

}  // namespace pph
