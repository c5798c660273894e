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
// 4 (l / 16) + r of column l % 16 of the result.  Tile 128 x 128 per
// workgroup of 4 waves (2 x 2, each 64 x 64 = 4 x 4 MFMA blocks), K in steps
// of 32; the next step's global loads are in flight while the current one
// is multiplied.
#include "common.hpp"
#include "kernels.hpp"

namespace pph {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int BT = 128;        // tile rows and columns
constexpr int BK = 32;         // K per step (one MFMA)
constexpr int BKP = BK + 8;    // LDS row length in BF16 (16 bytes of padding)

struct Bf16Gemm {
  // op(A)(m, k) = A[m * sam + k * sak], conjugated when ca; op(B)(k, n) =
  // B[k * sbk + n * sbn], conjugated when cb
  const cplx<float>* A;
  const cplx<float>* B;
  cplx<float>* C;
  int64_t sam, sak, sbk, sbn, strideA, strideB, ldc, strideC;
  int64_t M, N, K;
  int ca, cb;
};

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1)))
k_cgemm_bf16(Bf16Gemm p) {
  // three planes (re, im, re + im) of the A and B tiles, row-major in k
  __shared__ __align__(16) __bf16 As[3][BT][BKP];
  __shared__ __align__(16) __bf16 Bs[3][BT][BKP];
  const int tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
  const int64_t b = blockIdx.z;
  const int64_t m0 = (int64_t)blockIdx.x * BT, n0 = (int64_t)blockIdx.y * BT;
  const cplx<float>* A = p.A + b * p.strideA;
  const cplx<float>* B = p.B + b * p.strideB;
  // staging: work item w = tid + 256 h (h = 0, 1) is row (m or n) w % 128,
  // k group w / 128 (8 consecutive k)
  cplx<float> ra[2][8], rb[2][8];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int w = tid + 256 * h, r = w % BT, kg = w / BT;
      const int64_t m = m0 + r, nn = n0 + r;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int64_t k = k0 + 8 * kg + t;
        const bool oka = m < p.M && k < p.K, okb = nn < p.N && k < p.K;
        ra[h][t] = oka ? A[m * p.sam + k * p.sak] : cplx<float>{0.0f, 0.0f};
        rb[h][t] = okb ? B[k * p.sbk + nn * p.sbn] : cplx<float>{0.0f, 0.0f};
      }
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int w = tid + 256 * h, r = w % BT, kg = w / BT;
      bf16x8 ar, ai, as, br, bi, bs;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const float axr = ra[h][t].x, axi = p.ca ? -ra[h][t].y : ra[h][t].y;
        const float bxr = rb[h][t].x, bxi = p.cb ? -rb[h][t].y : rb[h][t].y;
        ar[t] = (__bf16)axr;
        ai[t] = (__bf16)axi;
        as[t] = (__bf16)(axr + axi);
        br[t] = (__bf16)bxr;
        bi[t] = (__bf16)bxi;
        bs[t] = (__bf16)(bxr + bxi);
      }
      *reinterpret_cast<bf16x8*>(&As[0][r][8 * kg]) = ar;
      *reinterpret_cast<bf16x8*>(&As[1][r][8 * kg]) = ai;
      *reinterpret_cast<bf16x8*>(&As[2][r][8 * kg]) = as;
      *reinterpret_cast<bf16x8*>(&Bs[0][r][8 * kg]) = br;
      *reinterpret_cast<bf16x8*>(&Bs[1][r][8 * kg]) = bi;
      *reinterpret_cast<bf16x8*>(&Bs[2][r][8 * kg]) = bs;
    }
  };
  f4 acc[3][4][4];
#pragma unroll
  for (int s = 0; s < 3; ++s)
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[s][i][j] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  const int wr = (wave / 2) * 64, wc = (wave % 2) * 64;
  const int lr = lane % 16, lk = 8 * (lane / 16);
  load(0);
  for (int64_t k0 = 0; k0 < p.K; k0 += BK) {
    store();
    __syncthreads();
    if (k0 + BK < p.K) load(k0 + BK);
    bf16x8 bo[3][4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int s = 0; s < 3; ++s) bo[s][j] = *reinterpret_cast<const bf16x8*>(&Bs[s][wc + 16 * j + lr][lk]);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      bf16x8 ao[3];
#pragma unroll
      for (int s = 0; s < 3; ++s) ao[s] = *reinterpret_cast<const bf16x8*>(&As[s][wr + 16 * i + lr][lk]);
#pragma unroll
      for (int j = 0; j < 4; ++j)
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
    for (int j = 0; j < 4; ++j) {
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
  p.M = g.m; p.N = g.n; p.K = g.k;
  const dim3 grid((unsigned)ceil_div(g.m, BT), (unsigned)ceil_div(g.n, BT), (unsigned)g.batch);
  hipLaunchKernelGGL(k_cgemm_bf16, grid, dim3(256), 0, s, p);
  PPH_CHECK_LAUNCH();
}

}  // namespace pph
