// Small epilogue kernels: unit conversion (diagonalize.F90:119 and
// write_output.F90:135), velocity scaling (write_output.F90:443-446),
// Rayleigh-quotient finish, precision conversion and zero fills.
#include "common.hpp"
#include "kernels.hpp"

namespace pph {

// f32-rounded literals as the Fortran source promotes them to f64
constexpr double C1 = 15.633301734924316;       // f32(15.633302)
constexpr double C2 = 33.356410980224609;       // f32(33.35641)
constexpr double C1SQ = 244.40011596679688;     // f32(f32(15.633302)**2)
constexpr double PI32 = 3.1415927410125732;     // f32(pi)

__global__ void k_fill_zero(uint4* p, int64_t n16) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = i; j < n16; j += stride) p[j] = make_uint4(0, 0, 0, 0);
}
__global__ void k_fill_zero_bytes(unsigned char* p, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

void launch_fill_zero(void* p, int64_t bytes, hipStream_t s) {
  if (bytes <= 0) return;
  const int64_t n16 = bytes / 16;
  if (n16 > 0) {
    const int64_t blocks = std::min<int64_t>(ceil_div(n16, 256), 65536);
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(256), 0, s, (uint4*)p, n16);
    PPH_CHECK_LAUNCH();
  }
  const int64_t rem = bytes - n16 * 16;
  if (rem > 0) {
    hipLaunchKernelGGL(k_fill_zero_bytes, dim3(1), dim3(16), 0, s, (unsigned char*)p + n16 * 16, rem);
    PPH_CHECK_LAUNCH();
  }
}

// omega = (sign(lam) sqrt|lam| * C1) * C2
__global__ void k_eval_to_cm(const double* lam, double* out, int64_t nb, int64_t m, int64_t ld) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t b = blockIdx.y;
  if (k >= m || b >= nb) return;
  const double l = lam[b * ld + k];
  const double s = sqrt(fabs(l)) * ((l < 0.0) ? -1.0 : 1.0);
  out[b * ld + k] = (s * C1) * C2;
}

void launch_eval_to_cm(const double* lam, double* out, int64_t nb, int64_t m, int64_t ld,
                       hipStream_t s) {
  if (m <= 0 || nb <= 0) return;
  hipLaunchKernelGGL(k_eval_to_cm, dim3((unsigned)ceil_div(m, 256), (unsigned)nb), dim3(256), 0, s,
                     lam, out, nb, m, ld);
  PPH_CHECK_LAUNCH();
}

// v[b][k][a] = ((((vp[b][a][k] / max(2|w|, 1e-3)) * C2) * C1SQ) * 100) * 2 * PI32, in place
// from the per-direction accumulators vp (3 x ldw per matrix) into v (ldv x 3 per matrix).
__global__ void k_vel_scale(double* vp, const double* om, int64_t nb, int64_t m, int64_t ldw,
                            int64_t ldv, double* vout) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t b = blockIdx.y;
  if (k >= m || b >= nb) return;
  const double den = fmax(2.0 * fabs(om[b * ldw + k]), 1e-3);
  for (int a = 0; a < 3; ++a) {
    const double d = vp[(b * 3 + a) * ldw + k];
    vout[b * ldv * 3 + k * 3 + a] = ((((d / den) * C2) * C1SQ) * 100.0) * 2.0 * PI32;
  }
}

void launch_vel_scale(double* v, const double* om_cm, int64_t nb, int64_t m, int64_t ldw,
                      int64_t ldv, hipStream_t s) {
  // v holds the accumulators on entry (3 x ldw per matrix) and is read only;
  // the scaled result goes to v + nb * 3 * ldw (ldv x 3 per matrix)
  if (m <= 0 || nb <= 0) return;
  hipLaunchKernelGGL(k_vel_scale, dim3((unsigned)ceil_div(m, 256), (unsigned)nb), dim3(256), 0, s,
                     v, om_cm, nb, m, ldw, ldv, v + nb * 3 * ldw);
  PPH_CHECK_LAUNCH();
}

template <typename T>
__global__ void __launch_bounds__(256) k_colnorm2(const cplx<T>* X, int64_t n, int64_t m,
                                                  int64_t ldx, int64_t strideX, double* out,
                                                  int64_t ldo) {
  __shared__ double sh[8];
  const int64_t k = blockIdx.x, b = blockIdx.y;
  if (k >= m) return;
  const cplx<T>* x = X + b * strideX + k * ldx;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const cplx<T> v = x[i];
    s += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[b * ldo + k] = s;
}

template <typename T>
void launch_colnorm2(const cplx<T>* X, int64_t n, int64_t m, int64_t ldx, int64_t strideX,
                     int64_t batch, double* out, int64_t ldo, hipStream_t s) {
  if (m <= 0 || batch <= 0) return;
  hipLaunchKernelGGL(k_colnorm2<T>, dim3((unsigned)m, (unsigned)batch), dim3(256), 0, s, X, n, m,
                     ldx, strideX, out, ldo);
  PPH_CHECK_LAUNCH();
}
template void launch_colnorm2<double>(const cplx<double>*, int64_t, int64_t, int64_t, int64_t,
                                      int64_t, double*, int64_t, hipStream_t);
template void launch_colnorm2<float>(const cplx<float>*, int64_t, int64_t, int64_t, int64_t,
                                     int64_t, double*, int64_t, hipStream_t);

__global__ void k_rq_finish(const double* num, const double* den, double* lam, int64_t nb,
                            int64_t m, int64_t ld) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t b = blockIdx.y;
  if (k >= m || b >= nb) return;
  lam[b * ld + k] = num[b * ld + k] / den[b * ld + k];
}

void launch_rq_finish(const double* num, const double* den, double* lam, int64_t nb, int64_t m,
                      int64_t ld, hipStream_t s) {
  if (m <= 0 || nb <= 0) return;
  hipLaunchKernelGGL(k_rq_finish, dim3((unsigned)ceil_div(m, 256), (unsigned)nb), dim3(256), 0, s,
                     num, den, lam, nb, m, ld);
  PPH_CHECK_LAUNCH();
}

// ---------------------------------------------------------------- column helpers
// Element (i, c) of matrix b sits at b * stride + c * ld + i; columns c < k.

__global__ void __launch_bounds__(256) k_coldot_re(const cplx<double>* A, const cplx<double>* Bm,
                                                   int64_t n, int64_t ld, int64_t stride,
                                                   double* out, int64_t ldo) {
  __shared__ double sh[8];
  const int64_t c = blockIdx.x, b = blockIdx.y;
  const cplx<double>* a = A + b * stride + c * ld;
  const cplx<double>* x = Bm + b * stride + c * ld;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += a[i].x * x[i].x + a[i].y * x[i].y;
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[b * ldo + c] = s;
}

void launch_coldot_re(const cplx<double>* A, const cplx<double>* Bm, int64_t n, int64_t k,
                      int64_t ld, int64_t stride, int64_t batch, double* out, int64_t ldo,
                      hipStream_t s) {
  if (k <= 0 || batch <= 0) return;
  hipLaunchKernelGGL(k_coldot_re, dim3((unsigned)k, (unsigned)batch), dim3(256), 0, s, A, Bm, n, ld,
                     stride, out, ldo);
  PPH_CHECK_LAUNCH();
}

__global__ void k_axpy_cols(cplx<double>* Y, const cplx<double>* X, const double* sc, double a,
                            int64_t n, int64_t ld, int64_t stride, int64_t lds) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t c = blockIdx.y, b = blockIdx.z;
  if (i >= n) return;
  const double f = a * sc[b * lds + c];
  const int64_t o = b * stride + c * ld + i;
  Y[o] = cplx<double>{Y[o].x + f * X[o].x, Y[o].y + f * X[o].y};
}

void launch_axpy_cols(cplx<double>* Y, const cplx<double>* X, const double* sc, double a, int64_t n,
                      int64_t k, int64_t ld, int64_t stride, int64_t lds, int64_t batch,
                      hipStream_t s) {
  if (n <= 0 || k <= 0 || batch <= 0) return;
  hipLaunchKernelGGL(k_axpy_cols, dim3((unsigned)ceil_div(n, 256), (unsigned)k, (unsigned)batch),
                     dim3(256), 0, s, Y, X, sc, a, n, ld, stride, lds);
  PPH_CHECK_LAUNCH();
}

__global__ void k_scale_rows(cplx<double>* T, const double* w, int64_t m, int64_t ld, int64_t stride,
                             int64_t ldw) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t c = blockIdx.y, b = blockIdx.z;
  if (j >= m) return;
  const double f = w[b * ldw + j];
  const int64_t o = b * stride + c * ld + j;
  // rows with weight 0 are cleared, whatever they hold
  T[o] = (f == 0.0) ? cplx<double>{0.0, 0.0} : cplx<double>{f * T[o].x, f * T[o].y};
}

void launch_scale_rows(cplx<double>* T, const double* w, int64_t m, int64_t k, int64_t ld,
                       int64_t stride, int64_t ldw, int64_t batch, hipStream_t s) {
  if (m <= 0 || k <= 0 || batch <= 0) return;
  hipLaunchKernelGGL(k_scale_rows, dim3((unsigned)ceil_div(m, 256), (unsigned)k, (unsigned)batch),
                     dim3(256), 0, s, T, w, m, ld, stride, ldw);
  PPH_CHECK_LAUNCH();
}

// conjugate-gradient step per column: al = rho / pq, z += al p, r -= al q;
// a column whose rho or pq is not positive (converged, or empty) stands still
__global__ void k_cg_update(cplx<double>* z, cplx<double>* r, const cplx<double>* p,
                            const cplx<double>* q, const double* rho, const double* pq, int64_t n,
                            int64_t ld, int64_t stride, int64_t lds) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t c = blockIdx.y, b = blockIdx.z;
  if (i >= n) return;
  const double rh = rho[b * lds + c], d = pq[b * lds + c];
  const double al = (rh > 0.0 && d > 0.0) ? rh / d : 0.0;
  const int64_t o = b * stride + c * ld + i;
  z[o] = cplx<double>{z[o].x + al * p[o].x, z[o].y + al * p[o].y};
  r[o] = cplx<double>{r[o].x - al * q[o].x, r[o].y - al * q[o].y};
}

// new direction per column: p = r + (rho_new / rho_old) p
__global__ void k_cg_direction(cplx<double>* p, const cplx<double>* r, const double* rho_new,
                               const double* rho_old, int64_t n, int64_t ld, int64_t stride,
                               int64_t lds) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t c = blockIdx.y, b = blockIdx.z;
  if (i >= n) return;
  const double ro = rho_old[b * lds + c];
  const double be = (ro > 0.0) ? rho_new[b * lds + c] / ro : 0.0;
  const int64_t o = b * stride + c * ld + i;
  p[o] = cplx<double>{r[o].x + be * p[o].x, r[o].y + be * p[o].y};
}

void launch_cg_update(cplx<double>* z, cplx<double>* r, const cplx<double>* p,
                      const cplx<double>* q, const double* rho, const double* pq, int64_t n,
                      int64_t k, int64_t ld, int64_t stride, int64_t lds, int64_t batch,
                      hipStream_t s) {
  if (n <= 0 || k <= 0 || batch <= 0) return;
  hipLaunchKernelGGL(k_cg_update, dim3((unsigned)ceil_div(n, 256), (unsigned)k, (unsigned)batch),
                     dim3(256), 0, s, z, r, p, q, rho, pq, n, ld, stride, lds);
  PPH_CHECK_LAUNCH();
}

void launch_cg_direction(cplx<double>* p, const cplx<double>* r, const double* rho_new,
                         const double* rho_old, int64_t n, int64_t k, int64_t ld, int64_t stride,
                         int64_t lds, int64_t batch, hipStream_t s) {
  if (n <= 0 || k <= 0 || batch <= 0) return;
  hipLaunchKernelGGL(k_cg_direction, dim3((unsigned)ceil_div(n, 256), (unsigned)k, (unsigned)batch),
                     dim3(256), 0, s, p, r, rho_new, rho_old, n, ld, stride, lds);
  PPH_CHECK_LAUNCH();
}

template <typename Ti, typename To>
__global__ void k_convert(const cplx<Ti>* in, cplx<To>* out, int64_t count) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = i; j < count; j += stride) {
    const cplx<Ti> v = in[j];
    out[j] = cplx<To>{(To)v.x, (To)v.y};
  }
}

template <typename Ti, typename To>
void launch_convert(const cplx<Ti>* in, cplx<To>* out, int64_t count, hipStream_t s) {
  if (count <= 0) return;
  const int64_t blocks = std::min<int64_t>(ceil_div(count, 256), 65536);
  hipLaunchKernelGGL((k_convert<Ti, To>), dim3((unsigned)blocks), dim3(256), 0, s, in, out, count);
  PPH_CHECK_LAUNCH();
}
template void launch_convert<float, double>(const cplx<float>*, cplx<double>*, int64_t, hipStream_t);
template void launch_convert<double, float>(const cplx<double>*, cplx<float>*, int64_t, hipStream_t);

// Batch outputs into the caller's device arrays in one launch per array:
// dst[b * dst_ld + i] = src[off[b] + i] for i < cnt[b] (doubles; off and cnt
// are B entries each), instead of one copy per q-point.
__global__ void __launch_bounds__(256) k_gather_rows(const double* __restrict__ src,
                                                     const int64_t* __restrict__ off,
                                                     const int64_t* __restrict__ cnt,
                                                     double* __restrict__ dst, int64_t dst_ld) {
  const int64_t b = blockIdx.y;
  const int64_t c = cnt[b];
  const double* sp = src + off[b];
  double* dp = dst + b * dst_ld;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < c;
       i += (int64_t)gridDim.x * blockDim.x)
    dp[i] = sp[i];
}

void launch_gather_rows(const double* src, const int64_t* off, const int64_t* cnt, double* dst,
                        int64_t dst_ld, int64_t max_cnt, int64_t B, hipStream_t s) {
  if (B <= 0 || max_cnt <= 0) return;
  const int64_t bx = std::min<int64_t>(ceil_div(max_cnt, 256), 1024);
  hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)bx, (unsigned)B), dim3(256), 0, s, src, off, cnt,
                     dst, dst_ld);
  PPH_CHECK_LAUNCH();
}

// info[b] |= flags[b] (per-matrix flags produced on the device)
__global__ void k_or_flags(int32_t* info, const int* flags, int64_t B) {
  const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b < B) info[b] |= flags[b];
}

void launch_or_flags(int32_t* info, const int* flags, int64_t B, hipStream_t s) {
  if (B <= 0) return;
  hipLaunchKernelGGL(k_or_flags, dim3((unsigned)ceil_div(B, 256)), dim3(256), 0, s, info, flags, B);
  PPH_CHECK_LAUNCH();
}

}  // namespace pph
