// Force-constant symmetrisation on the GPU: the iterations of the reference's
// symmetrize_fc (Force_Constant_Generation/get_force_constant.py:298-348) on
// the (N, N, 3, 3) array, C order, fc[i][j][a][b] at ((i N + j) 3 + a) 3 + b.
//
// One iteration = acoustic_sum_rule (:320-331): fc[i, :, a, b] -= its mean
// over the second atom, for every i, a, b; then inversion_symmetry
// (:334-348): fc[i, j, a, b] = fc[j, i, b, a] = their mean.  On the 3N x 3N
// matrix M[3i+a, 3j+b] = fc[i, j, a, b] the second step is M = (M + M^T)/2;
// the reference visits every unordered pair once with the values of the
// previous step, so the mean of each pair is exact and order-free.
#include "common.hpp"
#include "kernels.hpp"

namespace pph {

// one workgroup per atom i: row i is 9N contiguous doubles; thread t owns
// slot ab = t % 9 of atoms j = t / 9, t / 9 + 28, ...
constexpr int kAsrJ = 28;          // atoms per pass (28 x 9 = 252 threads)

__global__ void __launch_bounds__(256) k_fc_asr(double* fc, int64_t N) {
  __shared__ double part[9][kAsrJ];
  __shared__ double mean[9];
  double* row = fc + (int64_t)blockIdx.x * 9 * N;
  const int t = threadIdx.x;
  const int ab = t % 9, jj = t / 9;
  double s = 0.0;
  if (jj < kAsrJ)
    for (int64_t j = jj; j < N; j += kAsrJ) s += row[j * 9 + ab];
  if (jj < kAsrJ) part[ab][jj] = s;
  __syncthreads();
  if (t < 9) {
    double tot = 0.0;
    for (int k = 0; k < kAsrJ; ++k) tot += part[t][k];
    mean[t] = tot / (double)N;                 // np.mean = sum / N
  }
  __syncthreads();
  if (jj < kAsrJ) {
    const double mu = mean[ab];
    for (int64_t j = jj; j < N; j += kAsrJ) row[j * 9 + ab] -= mu;
  }
}

// 16 x 16 atom tiles (I, J), I <= J: both tiles staged in LDS, then every
// element and its mirror get (fc[i,j,a,b] + fc[j,i,b,a]) / 2
constexpr int kSymT = 16;

__global__ void __launch_bounds__(256) k_fc_sym(double* fc, int64_t N) {
  __shared__ double A[kSymT][kSymT * 9 + 1];   // fc[I0+i][J0+j][a][b] at A[i][j*9+ab]
  __shared__ double Bt[kSymT][kSymT * 9 + 1];  // fc[J0+j][I0+i][a][b] at Bt[j][i*9+ab]
  const int64_t I0 = (int64_t)blockIdx.x * kSymT, J0 = (int64_t)blockIdx.y * kSymT;
  if (I0 > J0) return;
  const int ni = (int)min((int64_t)kSymT, N - I0), nj = (int)min((int64_t)kSymT, N - J0);
  const int w = kSymT * 9;
  for (int e = threadIdx.x; e < kSymT * w; e += blockDim.x) {
    const int r = e / w, c = e % w;                   // c = j * 9 + ab
    if (r < ni && c / 9 < nj) A[r][c] = fc[((I0 + r) * N + J0) * 9 + c];
    if (r < nj && c / 9 < ni) Bt[r][c] = fc[((J0 + r) * N + I0) * 9 + c];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kSymT * w; e += blockDim.x) {
    const int i = e / w, c = e % w, j = c / 9, ab = c % 9, a = ab / 3, b = ab % 3;
    if (i >= ni || j >= nj) continue;
    const double m = (A[i][c] + Bt[j][i * 9 + b * 3 + a]) / 2;
    fc[((I0 + i) * N + J0 + j) * 9 + ab] = m;
    fc[((J0 + j) * N + I0 + i) * 9 + b * 3 + a] = m;
  }
}

void launch_fc_symmetrize(double* fc, int64_t N, int64_t iterations, hipStream_t s) {
  if (N <= 0) return;
  const unsigned nt = (unsigned)ceil_div(N, kSymT);
  for (int64_t it = 0; it < iterations; ++it) {
    hipLaunchKernelGGL(k_fc_asr, dim3((unsigned)N), dim3(256), 0, s, fc, N);
    PPH_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_fc_sym, dim3(nt, nt), dim3(256), 0, s, fc, N);
    PPH_CHECK_LAUNCH();
  }
}

}  // namespace pph
