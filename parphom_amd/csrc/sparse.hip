// Hermitian quadratic forms Re(z_i^H Op z_i) for every computed mode i, with
// Op = D(q) (the FP64 Rayleigh-quotient numerators) or dD/dq_{x,y,z} (the
// group-velocity accumulators of write_output.F90:410-446), from the nonzero
// 3x3 blocks of the operators only.  Zero blocks of the force constants give
// zero blocks of D and dD, so these sums equal the dense products' up to
// summation order; the dense path (gemm.hip, Hermitian-dot epilogue) is used
// when the blocks are not sparse.
//
// A block is a phase sum p times a real block Phi~ (per atom pair), so for a
// block (a, b) and mode z the forms of all operators share
//   c = z_a^H (Phi~ z_b)          (18 + 12 real FMAs)
// and each operator adds Re(p c).  Only blocks with a >= b are listed: an
// off-diagonal block stands for itself and its Hermitian partner (weight 2),
// a diagonal block (real symmetric Phi~, p real at q-independent zero
// separation) is taken whole, which equals the dense Hermitian dot's lower
// triangle plus Re(diagonal).
#include "common.hpp"
#include "kernels.hpp"

namespace pph {

constexpr int kDotAtoms = 16;      // atom rows per workgroup

// grid: x = 256 modes, y = kDotAtoms atom rows, z = q-point.  Thread = mode i,
// vectors from the mode-contiguous copy Zt (row r of matrix b at
// b * strideT + r * ldt).  Each workgroup stores its sums as partials
// part[((b * NOPS + o) * gridDim.y + blockIdx.y) * M + i]; k_dots_reduce adds
// them over the atom chunks in order.
template <typename Tz, int NOPS>
__global__ void __launch_bounds__(256) k_sparse_dots(const int64_t* __restrict__ rowptr,
                                                     const int32_t* __restrict__ bcol, int64_t nblk,
                                                     const double* __restrict__ phit,
                                                     const cplx<double>* __restrict__ bph, int op0,
                                                     const cplx<Tz>* __restrict__ Zt, int64_t ldt,
                                                     int64_t strideT, int64_t N, int64_t M,
                                                     double* __restrict__ part) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t b = blockIdx.z;
  if (i >= M) return;
  const cplx<Tz>* Z = Zt + b * strideT + i;
  const cplx<double>* P = bph + b * nblk * 4 + op0;
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
      double wr[3], wi[3];
      if (c < a) {
#pragma unroll
        for (int be = 0; be < 3; ++be) {
          const cplx<Tz> v = Z[(3 * c + be) * ldt];
          wr[be] = (double)v.x;
          wi[be] = (double)v.y;
        }
      } else {
#pragma unroll
        for (int be = 0; be < 3; ++be) { wr[be] = zr[be]; wi[be] = zi[be]; }
      }
      const double* F = phit + k * 9;
      double cr = 0.0, ci = 0.0;           // z_a^H (Phi~ w)
#pragma unroll
      for (int al = 0; al < 3; ++al) {
        const double ur = F[3 * al] * wr[0] + F[3 * al + 1] * wr[1] + F[3 * al + 2] * wr[2];
        const double ui = F[3 * al] * wi[0] + F[3 * al + 1] * wi[1] + F[3 * al + 2] * wi[2];
        cr += zr[al] * ur + zi[al] * ui;
        ci += zr[al] * ui - zi[al] * ur;
      }
      const double w = (c < a) ? 2.0 : 1.0;
#pragma unroll
      for (int o = 0; o < NOPS; ++o) {
        const cplx<double> p = P[k * 4 + o];
        acc[o] += w * (p.x * cr - p.y * ci);
      }
    }
  }
#pragma unroll
  for (int o = 0; o < NOPS; ++o)
    part[((b * NOPS + o) * (int64_t)gridDim.y + blockIdx.y) * M + i] = acc[o];
}

// out(op) += sum over the atom chunks of the partials, in order
__global__ void __launch_bounds__(256) k_dots_reduce(const double* __restrict__ part, int64_t ny,
                                                     int64_t M, int op0, int nops,
                                                     double* __restrict__ out0, int64_t out0_b,
                                                     double* __restrict__ out1, int64_t out1_op,
                                                     int64_t out1_b) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int o = blockIdx.y;
  const int64_t b = blockIdx.z;
  if (i >= M) return;
  const double* P = part + (b * nops + o) * ny * M + i;
  double acc = 0.0;
  for (int64_t y = 0; y < ny; ++y) acc += P[y * M];
  const int op = op0 + o;
  double* dst = (op == 0) ? out0 + b * out0_b + i : out1 + b * out1_b + (op - 1) * out1_op + i;
  *dst += acc;
}

int64_t sparse_dots_ws_bytes(int64_t N, int64_t M, int64_t batch, int nops) {
  return ceil_div(N, kDotAtoms) * M * batch * nops * (int64_t)sizeof(double);
}

template <typename Tz>
void launch_sparse_dots(const int64_t* rowptr, const int32_t* bcol, int64_t nblk,
                        const double* phit, const cplx<double>* bph, int op0, int nops,
                        const cplx<Tz>* Zt, int64_t ldt, int64_t strideT, int64_t N, int64_t M,
                        int64_t batch, double* out0, int64_t out0_b, double* out1,
                        int64_t out1_op, int64_t out1_b, double* ws, hipStream_t s) {
  if (M <= 0 || batch <= 0 || N <= 0) return;
  if (ws == nullptr) throw std::runtime_error("launch_sparse_dots: missing workspace");
  if (op0 < 0 || op0 + nops > 4) throw std::runtime_error("launch_sparse_dots: operators out of range");
  if ((op0 == 0 && !out0) || (op0 + nops > 1 && !out1))
    throw std::runtime_error("launch_sparse_dots: missing output");
  dim3 grid((unsigned)ceil_div(M, 256), (unsigned)ceil_div(N, kDotAtoms), (unsigned)batch);
#define PPH_DOTS(K)                                                                          \
  hipLaunchKernelGGL((k_sparse_dots<Tz, K>), grid, dim3(256), 0, s, rowptr, bcol, nblk, phit, \
                     bph, op0, Zt, ldt, strideT, N, M, ws)
  if (nops == 1) PPH_DOTS(1);
  else if (nops == 3) PPH_DOTS(3);
  else if (nops == 4) PPH_DOTS(4);
  else throw std::runtime_error("launch_sparse_dots: nops must be 1, 3 or 4");
#undef PPH_DOTS
  PPH_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_dots_reduce, dim3(grid.x, (unsigned)nops, (unsigned)batch), dim3(256), 0, s, ws,
                     (int64_t)grid.y, M, op0, nops, out0, out0_b, out1, out1_op, out1_b);
  PPH_CHECK_LAUNCH();
}

template void launch_sparse_dots<float>(const int64_t*, const int32_t*, int64_t, const double*,
                                        const cplx<double>*, int, int, const cplx<float>*, int64_t,
                                        int64_t, int64_t, int64_t, int64_t, double*, int64_t,
                                        double*, int64_t, int64_t, double*, hipStream_t);
template void launch_sparse_dots<double>(const int64_t*, const int32_t*, int64_t, const double*,
                                         const cplx<double>*, int, int, const cplx<double>*,
                                         int64_t, int64_t, int64_t, int64_t, int64_t, double*,
                                         int64_t, double*, int64_t, int64_t, double*, hipStream_t);

// Y = D X for k columns of X (n x k per matrix, column-major, ld n): thread =
// (atom row a, column c).  Row a of D is the listed blocks (a, b), b <= a,
// and the Hermitian partners of the blocks (c', a), c' > a, found through the
// transposed lists (tptr, trow, tblk): D_{a c'} = conj(p) Phi~^T.
__global__ void __launch_bounds__(256) k_sparse_spmm(const int64_t* __restrict__ rowptr,
                                                     const int32_t* __restrict__ bcol,
                                                     const int64_t* __restrict__ tptr,
                                                     const int32_t* __restrict__ trow,
                                                     const int64_t* __restrict__ tblk, int64_t nblk,
                                                     const double* __restrict__ phit,
                                                     const cplx<double>* __restrict__ bph,
                                                     const cplx<double>* __restrict__ X,
                                                     int64_t strideX, cplx<double>* __restrict__ Y,
                                                     int64_t strideY, int64_t N) {
  const int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t c = blockIdx.y, b = blockIdx.z;
  if (a >= N) return;
  const int64_t n = 3 * N;
  const cplx<double>* x = X + b * strideX + c * n;
  const cplx<double>* P = bph + b * nblk * 4;
  double yr[3] = {0, 0, 0}, yi[3] = {0, 0, 0};
  // accumulate p (F x_j) (transposed F when tr) into y
  auto add = [&](const double* F, cplx<double> p, int64_t j, bool tr) {
    double ur[3], ui[3];
#pragma unroll
    for (int al = 0; al < 3; ++al) {
      ur[al] = 0.0;
      ui[al] = 0.0;
#pragma unroll
      for (int be = 0; be < 3; ++be) {
        const double f = tr ? F[3 * be + al] : F[3 * al + be];
        const cplx<double> v = x[3 * j + be];
        ur[al] += f * v.x;
        ui[al] += f * v.y;
      }
    }
#pragma unroll
    for (int al = 0; al < 3; ++al) {
      yr[al] += p.x * ur[al] - p.y * ui[al];
      yi[al] += p.x * ui[al] + p.y * ur[al];
    }
  };
  for (int64_t k = rowptr[a]; k < rowptr[a + 1]; ++k) add(phit + k * 9, P[k * 4], bcol[k], false);
  for (int64_t t = tptr[a]; t < tptr[a + 1]; ++t) {
    const int64_t k = tblk[t];
    const cplx<double> p = P[k * 4];
    add(phit + k * 9, cplx<double>{p.x, -p.y}, trow[t], true);
  }
  cplx<double>* y = Y + b * strideY + c * n + 3 * a;
#pragma unroll
  for (int al = 0; al < 3; ++al) y[al] = cplx<double>{yr[al], yi[al]};
}

void launch_sparse_spmm(const int64_t* rowptr, const int32_t* bcol, const int64_t* tptr,
                        const int32_t* trow, const int64_t* tblk, int64_t nblk, const double* phit,
                        const cplx<double>* bph, const cplx<double>* X, int64_t strideX,
                        cplx<double>* Y, int64_t strideY, int64_t N, int64_t k, int64_t batch,
                        hipStream_t s) {
  if (N <= 0 || k <= 0 || batch <= 0) return;
  dim3 grid((unsigned)ceil_div(N, 256), (unsigned)k, (unsigned)batch);
  hipLaunchKernelGGL(k_sparse_spmm, grid, dim3(256), 0, s, rowptr, bcol, tptr, trow, tblk, nblk,
                     phit, bph, X, strideX, Y, strideY, N);
  PPH_CHECK_LAUNCH();
}

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
