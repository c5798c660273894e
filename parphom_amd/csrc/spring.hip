// Spring-model force constants on the GPU: the synthetic producer that
// stands in for the reference's finite-displacement LAMMPS producer
// (Force_Constant_Generation/get_force_constant.py:227-295), writing the
// dense (N, N, 3, 3) array parphom.x reads (force_read.F90:93-158) directly
// in device memory.
//
// Per listed pair (i, j) with minimum-image separation r (host list):
//   intralayer shell s:  Phi_ij = -(p_s rr + p_{3+s} (I - rr)) .* noise
//   interlayer:          Phi_ij = -(p_6 (exp(-(|r| - r0)/decay) rr)) .* noise
// with rr = r^ r^T, noise = 1 + amp u, u the symmetric counter-based numbers
// of the pair {i, j} (splitmix64 of seed * 0x9E3779B97F4A7C15 + (min N + max)
// * 8 + c for the six upper-triangle components, mapped to [-1, 1)); then the
// acoustic sum rule Phi_ii = -sum_{j != i} Phi_ij, summed over j ascending.
// The operations and their order are those of the CPU model
// (parphom_amd/synth.py), so the two agree to the rounding of exp.  The model
// is linear in p: with p = e_k the same kernels give dPhi/dp_k.
#include "common.hpp"
#include "kernels.hpp"

namespace pph {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) k_spring_pairs(int64_t N, int64_t npairs,
                                                      const int64_t* __restrict__ pi,
                                                      const int64_t* __restrict__ pj,
                                                      const double* __restrict__ rvec,
                                                      const int32_t* __restrict__ kind,
                                                      SpringParams sp, double* __restrict__ fc) {
#pragma clang fp contract(off)
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  const int64_t i = pi[p], j = pj[p];
  const double r0 = rvec[3 * p], r1 = rvec[3 * p + 1], r2 = rvec[3 * p + 2];
  const double dist = sqrt((r0 * r0 + r1 * r1) + r2 * r2);
  const double h[3] = {r0 / dist, r1 / dist, r2 / dist};
  const int kd = kind[p];
  const double kfac = (kd == 3) ? exp(-(dist - sp.inter_r0) / sp.inter_decay) : 0.0;
  const uint64_t key = (uint64_t)(i < j ? i : j) * (uint64_t)N + (uint64_t)(i < j ? j : i);
  const uint64_t base = sp.seed * 0x9E3779B97F4A7C15ull + key * 8ull;
  double u[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const uint64_t hsh = splitmix64(base + (uint64_t)c);
    u[c] = 2.0 * ((double)(hsh >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
  }
  // upper-triangle component index of (a, b)
  const int tri[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  double* out = fc + (i * N + j) * 9;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const double rr = h[a] * h[b];
      double acc;
      if (kd < 3) acc = sp.p[kd] * rr + sp.p[3 + kd] * ((a == b ? 1.0 : 0.0) - rr);
      else acc = sp.p[6] * (kfac * rr);
      const double noise = 1.0 + sp.amp * u[tri[a][b]];
      out[3 * a + b] = -acc * noise;
    }
}

// Phi_ii(a, b) = -sum_{j != i, ascending} Phi_ij(a, b): thread = (i, a, b)
__global__ void __launch_bounds__(256) k_spring_self(int64_t N, double* __restrict__ fc) {
#pragma clang fp contract(off)
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= N * 9) return;
  const int64_t i = t / 9, e = t % 9;
  const double* row = fc + i * N * 9 + e;
  double s = 0.0;
  for (int64_t j = 0; j < N; ++j)
    if (j != i) s += row[j * 9];
  fc[(i * N + i) * 9 + e] = -s;
}

void launch_spring_fc(int64_t N, int64_t npairs, const int64_t* pi, const int64_t* pj,
                      const double* rvec, const int32_t* kind, const SpringParams& sp, double* fc,
                      hipStream_t s) {
  launch_fill_zero(fc, N * N * 9 * (int64_t)sizeof(double), s);
  if (npairs > 0) {
    hipLaunchKernelGGL(k_spring_pairs, dim3((unsigned)ceil_div(npairs, 256)), dim3(256), 0, s, N,
                       npairs, pi, pj, rvec, kind, sp, fc);
    PPH_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_spring_self, dim3((unsigned)ceil_div(N * 9, 256)), dim3(256), 0, s, N, fc);
  PPH_CHECK_LAUNCH();
}

}  // namespace pph
