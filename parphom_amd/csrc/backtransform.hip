// Back-transformation of tridiagonal eigenvectors, Z = Q Z_T with
// Q = H_0 H_1 ... H_{n-2}, applied one panel of reflectors at a time in
// compact-WY form Q_p = I - V_p T_p V_p^H (the PZUNMTR stage of pzheevx):
//   Wt = V_p^H Z          (split-K)
//   VT = V_p T_p
//   Z  = Z - VT Wt
// all on the library's matrix-core GEMM.  T_p for every panel is formed up
// front from the Gram matrices V_p^H V_p.
#include "common.hpp"
#include "kernels.hpp"

namespace pph {

// Z(i, k) = Z_T(i, k) as complex, Z_T stored row-major (zt[i * ldzt + k]).
template <typename T>
__global__ void __launch_bounds__(256) k_zt_to_z(const double* __restrict__ zt, int64_t ldzt,
                                                 int64_t strideZt, cplx<T>* __restrict__ Z,
                                                 int64_t ldz, int64_t strideZ, int64_t n,
                                                 int64_t m) {
  __shared__ double tile[32][33];
  const int64_t b = blockIdx.z;
  const int64_t k0 = blockIdx.x * 32, i0 = blockIdx.y * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;   // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int64_t i = i0 + r, k = k0 + tx;
    tile[r][tx] = (i < n && k < m) ? zt[b * strideZt + i * ldzt + k] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int64_t k = k0 + r, i = i0 + tx;
    if (i < n && k < m) Z[b * strideZ + i + k * ldz] = cplx<T>{(T)tile[tx][r], (T)0};
  }
}

// zt[i ldzt + k] (FP64, real) -> complex64 (value, 0) in the same 8 bytes
__global__ void __launch_bounds__(256) k_zt_to_c64_rows(double* __restrict__ zt, int64_t ldzt,
                                                        int64_t strideZt, int64_t n, int64_t m) {
  const int64_t b = blockIdx.z, i = blockIdx.y;
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n || k >= m) return;
  double* p = zt + b * strideZt + i * ldzt + k;
  const double v = *p;
  *reinterpret_cast<cplx<float>*>(p) = cplx<float>{(float)v, 0.f};
}

// Z(i, k) = Zr(i, k): row-major to column-major through a 32 x 32 LDS tile
__global__ void __launch_bounds__(256) k_zr_to_z(const cplx<float>* __restrict__ Zr, int64_t ldr,
                                                 int64_t strideZr, cplx<float>* __restrict__ Z,
                                                 int64_t ldz, int64_t strideZ, int64_t n, int64_t m) {
  __shared__ cplx<float> tile[32][33];
  const int64_t b = blockIdx.z;
  const int64_t k0 = blockIdx.x * 32, i0 = blockIdx.y * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;   // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int64_t i = i0 + r, k = k0 + tx;
    tile[r][tx] = (i < n && k < m) ? Zr[b * strideZr + i * ldr + k] : cplx<float>{0, 0};
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int64_t k = k0 + r, i = i0 + tx;
    if (i < n && k < m) Z[b * strideZ + i + k * ldz] = tile[tx][r];
  }
}

void launch_zt_to_c64_rows(double* zt, int64_t ldzt, int64_t strideZt, int64_t n, int64_t m,
                           int64_t batch, hipStream_t s) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  dim3 grid((unsigned)ceil_div(m, 256), (unsigned)n, (unsigned)batch);
  hipLaunchKernelGGL(k_zt_to_c64_rows, grid, dim3(256), 0, s, zt, ldzt, strideZt, n, m);
  PPH_CHECK_LAUNCH();
}

void launch_zr_to_z(const cplx<float>* Zr, int64_t ldr, int64_t strideZr, cplx<float>* Z, int64_t ldz,
                    int64_t strideZ, int64_t n, int64_t m, int64_t batch, hipStream_t s) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  dim3 grid((unsigned)ceil_div(m, 32), (unsigned)ceil_div(n, 32), (unsigned)batch);
  hipLaunchKernelGGL(k_zr_to_z, grid, dim3(256), 0, s, Zr, ldr, strideZr, Z, ldz, strideZ, n, m);
  PPH_CHECK_LAUNCH();
}

// T_p from G_p = V_p^H V_p: T(k,k) = tau_k, T(0:k, k) = -tau_k T(0:k, 0:k) G(0:k, k).
// One workgroup (nb threads) per (block, matrix); entries beyond the block width are 0.
template <typename T>
__global__ void __launch_bounds__(256) k_form_t(const cplx<T>* __restrict__ G, cplx<T>* __restrict__ Tm,
                                               const cplx<T>* __restrict__ tau, int64_t strideTau,
                                               int64_t nrefl, int64_t nb, int64_t npanel,
                                               int64_t strideWs) {
  const int64_t p = blockIdx.x, b = blockIdx.y;
  const int64_t j0 = p * nb;
  const int64_t pw = min(nb, nrefl - j0);
  const cplx<T>* g = G + b * strideWs + p * nb * nb;
  cplx<T>* t = Tm + b * strideWs + p * nb * nb;
  const int i = threadIdx.x;
  for (int64_t c = 0; c < nb; ++c)
    if (i < nb) t[i + c * nb] = cplx<T>{0, 0};
  __syncthreads();
  for (int64_t k = 0; k < pw; ++k) {
    const cplx<T> tk = tau[b * strideTau + j0 + k];
    cplx<T> v{0, 0};
    if (i < k) {
      // (T(0:k,0:k) G(0:k,k))_i
      for (int64_t l = i; l < k; ++l) v = cadd(v, cmul(t[i + l * nb], g[l + k * nb]));
      v = cmul(cplx<T>{-tk.x, -tk.y}, v);
    }
    __syncthreads();
    if (i < k) t[i + k * nb] = v;
    if (i == k) t[k + k * nb] = tk;
    __syncthreads();
  }
}

template <typename T>
void launch_zt_to_z(const double* zt, int64_t ldzt, int64_t strideZt, cplx<T>* Z, int64_t ldz,
                    int64_t strideZ, int64_t n, int64_t m, int64_t batch, hipStream_t s) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  dim3 grid((unsigned)ceil_div(m, 32), (unsigned)ceil_div(n, 32), (unsigned)batch);
  hipLaunchKernelGGL(k_zt_to_z<T>, grid, dim3(256), 0, s, zt, ldzt, strideZt, Z, ldz, strideZ, n,
                     m);
  PPH_CHECK_LAUNCH();
}
template void launch_zt_to_z<float>(const double*, int64_t, int64_t, cplx<float>*, int64_t, int64_t,
                                    int64_t, int64_t, int64_t, hipStream_t);
template void launch_zt_to_z<double>(const double*, int64_t, int64_t, cplx<double>*, int64_t,
                                     int64_t, int64_t, int64_t, int64_t, hipStream_t);

template <typename T>
void back_transform(const BackWork<T>& w, hipStream_t s) {
  const int64_t n = w.n, nb = w.nb, M = w.m, B = w.batch;
  const int64_t sh = w.shift;
  if (M <= 0) return;
  const int64_t nrefl = std::max<int64_t>(0, n - sh);
  const int64_t npanel = ceil_div(nrefl, nb);
  // Z_T -> complex Z
  if (!w.z_ready) launch_zt_to_z<T>(w.zt, w.ldzt, w.strideZt, w.Z, w.ldz, w.strideZ, n, M, B, s);
  if (npanel == 0) return;
  // Gram matrices of all panels: G_p = V(:, p nb : p nb + nb)^H V(:, ...), all n rows
  cplx<T>* G = w.Wt;          // Wt workspace doubles as G storage before the panel loop
  {
    GemmParams<T> g{};
    g.opA = 'C'; g.opB = 'N';
    g.m = nb; g.n = nb; g.k = n;
    g.alpha = cplx<T>{1, 0}; g.beta = cplx<T>{0, 0};
    g.A = w.V; g.lda = n; g.strideA = w.strideV; g.sAin = nb * n;
    g.B = w.V; g.ldb = n; g.strideB = w.strideV; g.sBin = nb * n;
    g.C = G; g.ldc = nb; g.strideC = w.strideWs; g.sCin = nb * nb;
    g.batch = B; g.binner = npanel;
    launch_gemm<T>(g, s);
  }
  if (nb > 256) throw std::runtime_error("back_transform: block width must be <= 256");
  hipLaunchKernelGGL(k_form_t<T>, dim3((unsigned)npanel, (unsigned)B), dim3((unsigned)nb), 0, s, G, w.Tm,
                     w.tau, w.strideTau, nrefl, nb, npanel, w.strideWs);
  PPH_CHECK_LAUNCH();
  for (int64_t p = npanel - 1; p >= 0; --p) {
    const int64_t j0 = p * nb;
    const int64_t mr = n - j0 - sh;                // rows j0 + sh .. n-1
    const int64_t pw = std::min(nb, nrefl - j0);
    const int64_t kw = ceil_div(pw, 16) * 16;
    // Wt = V_p^H Z(j0 + sh:, :)   (kw x M)
    {
      GemmParams<T> g{};
      g.opA = 'C'; g.opB = 'N';
      g.m = kw; g.n = M; g.k = mr;
      g.alpha = cplx<T>{1, 0}; g.beta = cplx<T>{0, 0};
      g.A = w.V + (j0 + sh) + j0 * n; g.lda = n; g.strideA = w.strideV;
      g.B = w.Z + (j0 + sh); g.ldb = w.ldz; g.strideB = w.strideZ;
      g.C = w.Wt; g.ldc = nb; g.strideC = w.strideWs;
      g.batch = B;
      g.ksplit = back_ksplit(mr, M, B);
      g.ws = w.ws;
      if (g.ksplit > 1 && w.ws == nullptr) throw std::runtime_error("back_transform: split-K needs a workspace");
      launch_gemm<T>(g, s);
    }
    // VT = V_p T_p   (mr x kw)
    {
      GemmParams<T> g{};
      g.opA = 'N'; g.opB = 'N';
      g.m = mr; g.n = kw; g.k = kw;
      g.alpha = cplx<T>{1, 0}; g.beta = cplx<T>{0, 0};
      g.A = w.V + (j0 + sh) + j0 * n; g.lda = n; g.strideA = w.strideV;
      g.B = w.Tm + p * nb * nb; g.ldb = nb; g.strideB = w.strideWs;
      g.C = w.VT; g.ldc = n; g.strideC = w.strideWs;
      g.batch = B;
      launch_gemm<T>(g, s);
    }
    // Z(j0 + sh:, :) -= VT Wt
    {
      GemmParams<T> g{};
      g.opA = 'N'; g.opB = 'N';
      g.m = mr; g.n = M; g.k = kw;
      g.alpha = cplx<T>{-1, 0}; g.beta = cplx<T>{1, 0};
      g.A = w.VT; g.lda = n; g.strideA = w.strideWs;
      g.B = w.Wt; g.ldb = nb; g.strideB = w.strideWs;
      g.C = w.Z + (j0 + sh); g.ldc = w.ldz; g.strideC = w.strideZ;
      g.batch = B;
      launch_gemm<T>(g, s);
    }
  }
}

template void back_transform<double>(const BackWork<double>&, hipStream_t);
template void back_transform<float>(const BackWork<float>&, hipStream_t);

}  // namespace pph
