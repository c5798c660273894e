// Shared device/host helpers for the parphom HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <stdexcept>

namespace pph {

// ---------------------------------------------------------------- complex
template <typename T> struct alignas(2 * sizeof(T)) cplx { T x, y; };
template <typename T> struct vec2;
template <> struct vec2<double> { using type = double2; };
template <> struct vec2<float> { using type = float2; };

template <typename T>
__host__ __device__ __forceinline__ cplx<T> cmk(T r, T i) { return cplx<T>{r, i}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// conj(a) * b
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {
  return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cconj(cplx<T> a) { return {a.x, -a.y}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) { return {a.x * s, a.y * s}; }
template <typename T>
__host__ __device__ __forceinline__ T cabs2(cplx<T> a) { return a.x * a.x + a.y * a.y; }

// ---------------------------------------------------------------- reductions
constexpr int WAVE = 64;

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
  return v;
}

// Block-wide sum; `sh` must hold blockDim.x / WAVE entries.  All threads get the result.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  T r = 0;
  for (int i = 0; i < nw; ++i) r += sh[i];
  __syncthreads();
  return r;
}

template <typename T>
__device__ __forceinline__ T block_max(T v, T* sh) {
  v = wave_max(v);
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  T r = sh[0];
  for (int i = 1; i < nw; ++i) r = fmax(r, sh[i]);
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------- matrix cores
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct mfma;
// Operand layout of both 16x16x4 instructions: lane l supplies A[l % 16][l / 16]
// and B[l / 16][l % 16].  Result: lane l holds column l % 16; accumulator
// register r holds row 4 * (l / 16) + r for f32 but row 4 * r + l / 16 for f64.
template <> struct mfma<double> {
  typedef d4 acc;
  static __device__ __forceinline__ acc op(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * r + lane / 16; }
};
template <> struct mfma<float> {
  typedef f4 acc;
  static __device__ __forceinline__ acc op(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane / 16) + r; }
};

// ---------------------------------------------------------------- Householder reflector
// H = I - tau v v^H with v = (1, x / (alpha - beta)) and H^H (alpha, x) = (beta, 0):
// beta = -sign(Re alpha) sqrt(|alpha|^2 + |x|^2), tau = (beta - alpha) / beta
// (the ZLARFG formulas without the rescaling loop); x = 0 with a real alpha
// gives tau = 0.  sr, si: scale = 1 / (alpha - beta).
struct Reflector { double beta, tr, ti, sr, si; };

__device__ __forceinline__ Reflector make_reflector(double xn2, double ar, double ai) {
  Reflector r{ar, 0, 0, 0, 0};
  if (!(xn2 == 0.0 && ai == 0.0)) {
    const double nrm = sqrt(ar * ar + ai * ai + xn2);
    r.beta = (ar >= 0.0) ? -nrm : nrm;
    r.tr = (r.beta - ar) / r.beta;
    r.ti = -ai / r.beta;
    const double dr = ar - r.beta, di = ai;
    const double den = dr * dr + di * di;
    r.sr = dr / den;
    r.si = -di / den;
  }
  return r;
}

// ---------------------------------------------------------------- host errors
struct hip_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define PPH_HIP(call)                                                             \
  do {                                                                            \
    hipError_t _e = (call);                                                       \
    if (_e != hipSuccess)                                                         \
      throw ::pph::hip_error(std::string(#call) + ": " + hipGetErrorString(_e) + \
                             " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
  } while (0)

#define PPH_CHECK_LAUNCH() PPH_HIP(hipGetLastError())

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
__device__ __forceinline__ int64_t ceil_div_dev(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace pph
