// Hermitian quadratic forms Re(z_i^H Op z_i) for every computed mode i, with
// Op = D(q) (the FP64 Rayleigh-quotient numerators) or dD/dq_{x,y,z} (the
// group-velocity accumulators of write_output.F90:410-446), from the nonzero
// 3x3 blocks of the operators only.  Zero blocks of the force constants give
// zero blocks of D and dD, so these sums equal the dense products' up to
// summation order; the dense path (gemm.hip, Hermitian-dot epilogue) is used
// when the blocks are not sparse.
//
// Like the dense Hermitian dot, only the lower triangle is read:
//   z^H Op z = 2 Re(z^H Op' z), Op' = strict lower triangle + Re(diag) / 2,
// here per block (a, b), a >= b: weight 2 for a > b; inside a diagonal block
// weight 2 below its diagonal and Re(Op) |z|^2 on it.
#include "common.hpp"
#include "kernels.hpp"

namespace pph {

constexpr int kDotAtoms = 16;      // atom rows per workgroup

// grid: x = 256 modes, y = kDotAtoms atom rows, z = q-point.  Thread = mode i,
// vectors from the mode-contiguous copy Zt (row r of matrix b at
// b * strideT + r * ldt).  Block values bv as written by k_sparse_blocks.
template <typename Tz, int NOPS>
__global__ void __launch_bounds__(256) k_sparse_dots(const int64_t* __restrict__ rowptr,
                                                     const int32_t* __restrict__ bcol, int64_t nblk,
                                                     const cplx<double>* __restrict__ bv, int op0,
                                                     const cplx<Tz>* __restrict__ Zt, int64_t ldt,
                                                     int64_t strideT, int64_t N, int64_t M,
                                                     double* __restrict__ out, int64_t out_op,
                                                     int64_t out_b) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t b = blockIdx.z;
  if (i >= M) return;
  const cplx<Tz>* Z = Zt + b * strideT + i;
  double acc[NOPS];
#pragma unroll
  for (int o = 0; o < NOPS; ++o) acc[o] = 0.0;
  const int64_t a0 = (int64_t)blockIdx.y * kDotAtoms;
  const int64_t a1 = min(N, a0 + kDotAtoms);
  for (int64_t a = a0; a < a1; ++a) {
    double zr[3], zi[3];
#pragma unroll
    for (int al = 0; al < 3; ++al) {
      const cplx<Tz> v = Z[(3 * a + al) * ldt];
      zr[al] = (double)v.x;
      zi[al] = (double)v.y;
    }
    const int64_t k1 = rowptr[a + 1];
    for (int64_t k = rowptr[a]; k < k1; ++k) {
      const int64_t c = bcol[k];
      const cplx<double>* V = bv + (b * nblk + k) * 36 + 9 * op0;
      if (c < a) {
        double wr[3], wi[3];
#pragma unroll
        for (int be = 0; be < 3; ++be) {
          const cplx<Tz> v = Z[(3 * c + be) * ldt];
          wr[be] = (double)v.x;
          wi[be] = (double)v.y;
        }
#pragma unroll
        for (int o = 0; o < NOPS; ++o) {
          double s = 0.0;
#pragma unroll
          for (int al = 0; al < 3; ++al) {
            double tr = 0.0, ti = 0.0;       // (Op w)_alpha
#pragma unroll
            for (int be = 0; be < 3; ++be) {
              const cplx<double> e = V[9 * o + 3 * al + be];
              tr += e.x * wr[be] - e.y * wi[be];
              ti += e.x * wi[be] + e.y * wr[be];
            }
            s += zr[al] * tr + zi[al] * ti;   // Re(conj(z_alpha) (Op w)_alpha)
          }
          acc[o] += 2.0 * s;
        }
      } else {                               // diagonal block (c == a)
#pragma unroll
        for (int o = 0; o < NOPS; ++o) {
          double s = 0.0;
#pragma unroll
          for (int al = 0; al < 3; ++al) {
            s += V[9 * o + 4 * al].x * (zr[al] * zr[al] + zi[al] * zi[al]);
#pragma unroll
            for (int be = 0; be < al; ++be) {
              const cplx<double> e = V[9 * o + 3 * al + be];
              const double tr = e.x * zr[be] - e.y * zi[be];
              const double ti = e.x * zi[be] + e.y * zr[be];
              s += 2.0 * (zr[al] * tr + zi[al] * ti);
            }
          }
          acc[o] += s;
        }
      }
    }
  }
#pragma unroll
  for (int o = 0; o < NOPS; ++o) atomicAdd(out + b * out_b + o * out_op + i, acc[o]);
}

template <typename Tz>
void launch_sparse_dots(const int64_t* rowptr, const int32_t* bcol, int64_t nblk,
                        const cplx<double>* bv, int op0, int nops, const cplx<Tz>* Zt, int64_t ldt,
                        int64_t strideT, int64_t N, int64_t M, int64_t batch, double* out,
                        int64_t out_op, int64_t out_b, hipStream_t s) {
  if (M <= 0 || batch <= 0 || N <= 0) return;
  dim3 grid((unsigned)ceil_div(M, 256), (unsigned)ceil_div(N, kDotAtoms), (unsigned)batch);
  if (nops == 1)
    hipLaunchKernelGGL((k_sparse_dots<Tz, 1>), grid, dim3(256), 0, s, rowptr, bcol, nblk, bv, op0,
                       Zt, ldt, strideT, N, M, out, out_op, out_b);
  else if (nops == 3)
    hipLaunchKernelGGL((k_sparse_dots<Tz, 3>), grid, dim3(256), 0, s, rowptr, bcol, nblk, bv, op0,
                       Zt, ldt, strideT, N, M, out, out_op, out_b);
  else
    throw std::runtime_error("launch_sparse_dots: nops must be 1 or 3");
  PPH_CHECK_LAUNCH();
}

template void launch_sparse_dots<float>(const int64_t*, const int32_t*, int64_t, const cplx<double>*,
                                        int, int, const cplx<float>*, int64_t, int64_t, int64_t,
                                        int64_t, int64_t, double*, int64_t, int64_t, hipStream_t);
template void launch_sparse_dots<double>(const int64_t*, const int32_t*, int64_t,
                                         const cplx<double>*, int, int, const cplx<double>*, int64_t,
                                         int64_t, int64_t, int64_t, int64_t, double*, int64_t,
                                         int64_t, hipStream_t);

// Zt(r, i) = Z(r, i) for rows r < n and modes i < m: Z column-major (ldz),
// Zt mode-contiguous (row r at r * ldt).  32 x 32 tiles through LDS.
template <typename T>
__global__ void __launch_bounds__(256) k_transpose(const cplx<T>* __restrict__ Z, int64_t n,
                                                   int64_t m, int64_t ldz, int64_t strideZ,
                                                   cplx<T>* __restrict__ Zt, int64_t ldt,
                                                   int64_t strideT) {
  __shared__ cplx<T> tile[32][33];
  const int64_t b = blockIdx.z;
  const int64_t r0 = (int64_t)blockIdx.x * 32, i0 = (int64_t)blockIdx.y * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;   // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int64_t r = r0 + tx, i = i0 + k;
    if (r < n && i < m) tile[k][tx] = Z[b * strideZ + r + i * ldz];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int64_t r = r0 + k, i = i0 + tx;
    if (r < n && i < m) Zt[b * strideT + r * ldt + i] = tile[tx][k];
  }
}

template <typename T>
void launch_transpose(const cplx<T>* Z, int64_t n, int64_t m, int64_t ldz, int64_t strideZ,
                      cplx<T>* Zt, int64_t ldt, int64_t strideT, int64_t batch, hipStream_t s) {
  if (n <= 0 || m <= 0 || batch <= 0) return;
  dim3 grid((unsigned)ceil_div(n, 32), (unsigned)ceil_div(m, 32), (unsigned)batch);
  hipLaunchKernelGGL(k_transpose<T>, grid, dim3(256), 0, s, Z, n, m, ldz, strideZ, Zt, ldt, strideT);
  PPH_CHECK_LAUNCH();
}

template void launch_transpose<float>(const cplx<float>*, int64_t, int64_t, int64_t, int64_t,
                                      cplx<float>*, int64_t, int64_t, int64_t, hipStream_t);
template void launch_transpose<double>(const cplx<double>*, int64_t, int64_t, int64_t, int64_t,
                                       cplx<double>*, int64_t, int64_t, int64_t, hipStream_t);

}  // namespace pph
