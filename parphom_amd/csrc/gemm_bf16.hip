// Complex GEMM on the BF16 matrix cores, C = op(A) op(B) with complex64
// inputs and output and FP32 accumulation: the inputs are rounded to BF16
// (8-bit significand) as they are staged into LDS.  For products whose
// result is itself a small correction (the first step of the vector
// correction of the FP32 path, vecref.hip), where a relative error of about
// 2^-9 of the terms is enough and the BF16 rate (16 times the FP32 matrix
// rate on gfx950) pays.
//
// 3M scheme (three real products per complex step): P1 += Ar Br,
// P2 += Ai Bi, P3 += (Ar + Ai)(Br + Bi), Re = P1 - P2, Im = P3 - P1 - P2;
// the three planes Ar, Ai, Ar + Ai (and the same of B) are formed in FP32
// and rounded once.  v_mfma_f32_16x16x32_bf16: lane l supplies row l % 16 of
// A and column l % 16 of B at k = 8 (l / 16) .. + 7, and holds rows
// 4 (l / 16) + r of column l % 16 of the result.  Tile 128 x 64 per
// workgroup of 4 waves (2 x 2, each 64 x 32 = 4 x 2 MFMA blocks), K in steps
// of 32; the next step's global loads are in flight while the current one
// is multiplied.
//
// Tile order: the dispatcher hands workgroup i to XCD i mod 8, and each XCD
// has its own L2.  The kernel renumbers its workgroups so that XCD x takes a
// contiguous eighth of the (matrix, tile) list, walked in groups of 8 row
// tiles by all column tiles: the workgroups in flight on one XCD then share
// 8 row panels of A and a few column panels of B in that XCD's L2, instead
// of every XCD streaming every panel.
#include "common.hpp"
#include "kernels.hpp"

#include <type_traits>

namespace pph {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int BT = 128;        // tile rows and columns
constexpr int BK = 32;         // K per step (one MFMA)
constexpr int BKP = BK + 8;    // LDS row length in BF16 (16 bytes of padding)
constexpr int kBf16TileN = 64; // tile columns (64: two waves per SIMD)
constexpr int kXcds = 8;       // XCDs of an MI355X (round-robin workgroup dispatch)
constexpr int kGroupM = 8;     // row tiles per group of the tile order

struct Bf16Gemm {
  // op(A)(m, k) = A[m * sam + k * sak], conjugated when ca; op(B)(k, n) =
  // B[k * sbk + n * sbn], conjugated when cb
  const cplx<float>* A;
  const cplx<float>* B;
  cplx<float>* C;
  int64_t sam, sak, sbk, sbn, strideA, strideB, ldc, strideC;
  int64_t M, N, K, batch;
  int ca, cb;
};

// KA / KB: the operand is contiguous along k (op 'C' of A, op 'N' of B, with
// 16-byte aligned pairs): it is loaded two elements (16 bytes) per lane, 16
// lanes per 256-byte row segment of the tile; otherwise contiguous along its
// rows (columns): one element per lane, consecutive lanes on consecutive rows
// TN: tile columns, 128 (each wave 64 x 64, one wave per SIMD) or 64 (each
// wave 64 x 32: fewer registers, two waves per SIMD hide the global loads)
template <bool KA, bool KB, int TN>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TN == 128 ? 1 : 2)))
k_cgemm_bf16(Bf16Gemm p) {
  constexpr int NJ = TN / 32;          // 16-column MFMA blocks per wave
  // three planes (re, im, re + im) of the A and B tiles, row-major in k
  __shared__ __align__(16) __bf16 As[3][BT][BKP];
  __shared__ __align__(16) __bf16 Bs[3][TN][BKP];
  const int tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
  // (matrix, row tile, column tile) of this workgroup (tile order above)
  const int64_t gm = (p.M + BT - 1) / BT, gn = (p.N + TN - 1) / TN, per = gm * gn;
  const int64_t total = per * p.batch, whole = total / kXcds * kXcds;
  int64_t lid = blockIdx.x;
  if (lid < whole) lid = (lid % kXcds) * (total / kXcds) + lid / kXcds;
  const int64_t b = lid / per, rt = lid % per;
  const int64_t first = rt / (kGroupM * gn) * kGroupM, rows = min(gm - first, (int64_t)kGroupM);
  const int64_t m0 = (first + rt % (kGroupM * gn) % rows) * BT;
  const int64_t n0 = rt % (kGroupM * gn) / rows * TN;
  const cplx<float>* A = p.A + b * p.strideA;
  const cplx<float>* B = p.B + b * p.strideB;
  // 16 elements of each operand per thread and K step: k-contiguous, pair e
  // = tid + 256 h (h < 8) is row e / 16, k = 2 (e % 16) (+1); otherwise item
  // w = tid + 256 h (h < 2) is row w % 128, k = 8 (w / 128) + t
  cplx<float> ra[16], rb[TN / 8];
  // RT rows of the operand's tile, RT / 8 elements per thread: NH = RT / 16
  // pairs (k-contiguous) or NH / 4 items of 8 consecutive k
  auto load_op = [&](cplx<float>* reg, const cplx<float>* X, int64_t s_row, int64_t s_k, int64_t r0,
                     int64_t R, int64_t k0, bool kcont, auto rt) {
    constexpr int RT = decltype(rt)::value, NH = RT / 16;
    if (kcont) {
#pragma unroll
      for (int h = 0; h < NH; ++h) {
        const int e = tid + 256 * h, r = e / 16, kp = e % 16;
        const int64_t row = r0 + r, k = k0 + 2 * kp;
        if (row < R && k + 1 < p.K) {
          const float4 v = *reinterpret_cast<const float4*>(X + row * s_row + k);
          reg[2 * h] = cplx<float>{v.x, v.y};
          reg[2 * h + 1] = cplx<float>{v.z, v.w};
        } else {
          reg[2 * h] = (row < R && k < p.K) ? X[row * s_row + k] : cplx<float>{0.0f, 0.0f};
          reg[2 * h + 1] = cplx<float>{0.0f, 0.0f};
        }
      }
    } else {
#pragma unroll
      for (int h = 0; h < NH / 4; ++h) {
        const int w = tid + 256 * h, r = w % RT, kg = w / RT;
        const int64_t row = r0 + r;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const int64_t k = k0 + 8 * kg + t;
          reg[8 * h + t] = (row < R && k < p.K) ? X[row * s_row + k * s_k] : cplx<float>{0.0f, 0.0f};
        }
      }
    }
  };
  auto load = [&](int64_t k0) {
    load_op(ra, A, p.sam, p.sak, m0, p.M, k0, KA, std::integral_constant<int, BT>{});
    load_op(rb, B, p.sbn, p.sbk, n0, p.N, k0, KB, std::integral_constant<int, TN>{});
  };
  auto store_op = [&](const cplx<float>* reg, __bf16* S, int conj, bool kcont, auto rt) {
    constexpr int RT = decltype(rt)::value, NH = RT / 16;
    auto at = [&](int pl, int r, int k) { return S + ((size_t)pl * RT + r) * BKP + k; };
    if (kcont) {
      typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
#pragma unroll
      for (int h = 0; h < NH; ++h) {
        const int e = tid + 256 * h, r = e / 16, kp = e % 16;
        bf16x2 vr, vi, vs;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          const float xr = reg[2 * h + t].x, xi = conj ? -reg[2 * h + t].y : reg[2 * h + t].y;
          vr[t] = (__bf16)xr;
          vi[t] = (__bf16)xi;
          vs[t] = (__bf16)(xr + xi);
        }
        *reinterpret_cast<bf16x2*>(at(0, r, 2 * kp)) = vr;
        *reinterpret_cast<bf16x2*>(at(1, r, 2 * kp)) = vi;
        *reinterpret_cast<bf16x2*>(at(2, r, 2 * kp)) = vs;
      }
    } else {
#pragma unroll
      for (int h = 0; h < NH / 4; ++h) {
        const int w = tid + 256 * h, r = w % RT, kg = w / RT;
        bf16x8 vr, vi, vs;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const float xr = reg[8 * h + t].x, xi = conj ? -reg[8 * h + t].y : reg[8 * h + t].y;
          vr[t] = (__bf16)xr;
          vi[t] = (__bf16)xi;
          vs[t] = (__bf16)(xr + xi);
        }
        *reinterpret_cast<bf16x8*>(at(0, r, 8 * kg)) = vr;
        *reinterpret_cast<bf16x8*>(at(1, r, 8 * kg)) = vi;
        *reinterpret_cast<bf16x8*>(at(2, r, 8 * kg)) = vs;
      }
    }
  };
  auto store = [&]() {
    store_op(ra, &As[0][0][0], p.ca, KA, std::integral_constant<int, BT>{});
    store_op(rb, &Bs[0][0][0], p.cb, KB, std::integral_constant<int, TN>{});
  };
  f4 acc[3][4][NJ];
#pragma unroll
  for (int s = 0; s < 3; ++s)
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < NJ; ++j) acc[s][i][j] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  const int wr = (wave / 2) * 64, wc = (wave % 2) * (TN / 2);
  const int lr = lane % 16, lk = 8 * (lane / 16);
  load(0);
  for (int64_t k0 = 0; k0 < p.K; k0 += BK) {
    store();
    __syncthreads();
    if (k0 + BK < p.K) load(k0 + BK);
    bf16x8 bo[3][NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int s = 0; s < 3; ++s) bo[s][j] = *reinterpret_cast<const bf16x8*>(&Bs[s][wc + 16 * j + lr][lk]);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      bf16x8 ao[3];
#pragma unroll
      for (int s = 0; s < 3; ++s) ao[s] = *reinterpret_cast<const bf16x8*>(&As[s][wr + 16 * i + lr][lk]);
#pragma unroll
      for (int j = 0; j < NJ; ++j)
#pragma unroll
        for (int s = 0; s < 3; ++s)
          acc[s][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ao[s], bo[s][j], acc[s][i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  cplx<float>* C = p.C + b * p.strideC;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int64_t col = n0 + wc + 16 * j + lr;
      if (col >= p.N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wr + 16 * i + 4 * (lane / 16) + r;
        if (row >= p.M) continue;
        const float p1 = acc[0][i][j][r], p2 = acc[1][i][j][r], p3 = acc[2][i][j][r];
        C[row + col * p.ldc] = cplx<float>{p1 - p2, p3 - p1 - p2};
      }
    }
}

void launch_cgemm_bf16(const GemmParams<float>& g, hipStream_t s) {
  if (g.m <= 0 || g.n <= 0 || g.batch <= 0) return;
  if (g.A2 || g.dot_out || g.ksplit > 1 || g.tri || g.lower || g.binner > 1)
    throw std::runtime_error("launch_cgemm_bf16: one plain product only");
  if (g.alpha.x != 1.0f || g.alpha.y != 0.0f || g.beta.x != 0.0f || g.beta.y != 0.0f)
    throw std::runtime_error("launch_cgemm_bf16: alpha 1 and beta 0 only");
  Bf16Gemm p{};
  p.A = g.A; p.B = g.B; p.C = g.C;
  // op(A)(m, k): A(m, k) for 'N', conj(A(k, m)) for 'C' (column-major, lda)
  if (g.opA == 'N') { p.sam = 1; p.sak = g.lda; p.ca = 0; }
  else if (g.opA == 'C') { p.sam = g.lda; p.sak = 1; p.ca = 1; }
  else throw std::runtime_error("launch_cgemm_bf16: opA must be 'N' or 'C'");
  if (g.opB == 'N') { p.sbk = 1; p.sbn = g.ldb; p.cb = 0; }
  else if (g.opB == 'C') { p.sbk = g.ldb; p.sbn = 1; p.cb = 1; }
  else throw std::runtime_error("launch_cgemm_bf16: opB must be 'N' or 'C'");
  p.strideA = g.strideA; p.strideB = g.strideB; p.ldc = g.ldc; p.strideC = g.strideC;
  p.M = g.m; p.N = g.n; p.K = g.k; p.batch = g.batch;
  // the 16-byte pair loads need a unit k stride, even leading dimensions and
  // strides and 16-byte aligned bases
  auto pairs_ok = [](const void* base, int64_t ld, int64_t stride, int64_t sk) {
    return sk == 1 && ld % 2 == 0 && stride % 2 == 0 && (reinterpret_cast<uintptr_t>(base) & 15) == 0;
  };
  const bool ka = pairs_ok(p.A, p.sam, p.strideA, p.sak);
  const bool kb = pairs_ok(p.B, p.sbn, p.strideB, p.sbk);
  constexpr int TN = kBf16TileN;
  const int64_t tiles = ceil_div(g.m, BT) * ceil_div(g.n, TN) * g.batch;
  if (tiles > INT32_MAX) throw std::runtime_error("launch_cgemm_bf16: too many tiles");
  const dim3 grid((unsigned)tiles);
  if (ka && kb) hipLaunchKernelGGL((k_cgemm_bf16<true, true, TN>), grid, dim3(256), 0, s, p);
  else if (ka) hipLaunchKernelGGL((k_cgemm_bf16<true, false, TN>), grid, dim3(256), 0, s, p);
  else if (kb) hipLaunchKernelGGL((k_cgemm_bf16<false, true, TN>), grid, dim3(256), 0, s, p);
  else hipLaunchKernelGGL((k_cgemm_bf16<false, false, TN>), grid, dim3(256), 0, s, p);
  PPH_CHECK_LAUNCH();
}

}  // namespace pph
