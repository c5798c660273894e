// Phonon density of states on the GPU: linear triangulation and Gaussian /
// Lorentzian smearing, the integrals of the reference's utility layer
// (Utilities/pyphutil/fortran_routines/bz_integration.f90).
//
// Linear triangulation (bz_integration.f90:12-138).  For every energy E_i,
// triangle t and band b the reference sorts [E_i, (f, g) at the three
// vertices] with a 4-element exchange sort and evaluates one of four closed
// forms chosen by the position of E_i.  Outside [min f, max f] the result
// does not depend on E: (0, 0) below, (0, g0 + g1 + g2) above.  So one
// thread per (t, b) visits only the energies inside its window, evaluating
// the reference's sort and branches exactly, and accumulates with FP64
// atomics; the constant above the window goes into a difference array that
// a prefix sum turns into the number density.
//
// Smearing (bz_integration.f90:159-199).  dos_i = sum_q (sum_b g(E_i - f_qb))
// * w_q / natom.  Workgroups own 256 consecutive energies and a range of
// q-points; frequencies are staged through LDS in tiles that never cross a
// q-point.  Each q-point's eigenvalues are ascending, so a Gaussian tile more
// than 40 widths away from the workgroup's energies is skipped: every such
// term underflows to exactly 0 in the reference too (exp(-800)).  The
// reference's PI literal is a default real, so both shapes use float32 pi.
#include "common.hpp"
#include "kernels.hpp"

namespace pph {

struct Vtx { double e, f; };

// bz_integration.f90:49-64 on [E, v0, v1, v2]
__device__ __forceinline__ void sort4(Vtx (&a)[4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = i + 1; j < 4; ++j)
      if (a[i].e > a[j].e) {
        const Vtx t = a[i];
        a[i] = a[j];
        a[j] = t;
      }
}

// bz_integration.f90:67-134: the four position tests in the reference's
// order, each overwriting the previous result
__device__ __forceinline__ void integrate_triangle(const Vtx (&t)[4], double Ef, double& dos,
                                                   double& nd) {
#pragma clang fp contract(off)
  dos = 0.0;
  nd = 0.0;
  if (t[3].e == Ef) {
    dos = 0.0;
    nd = (t[0].f + t[1].f) + t[2].f;
  }
  if (t[2].e == Ef) {
    const double e0 = t[0].e, f0 = t[0].f, e1 = t[1].e, f1 = t[1].f, e2 = t[3].e, f2 = t[3].f;
    const double del02 = (e2 - Ef) / (e2 - e0);
    const double del12 = (e2 - Ef) / (e2 - e1);
    const double ddel02 = 1 / (e0 - e2);
    const double ddel12 = 1 / (e1 - e2);
    nd = (1 - del12 * (del02 * del02)) * f0 + (1 - del02 * (del12 * del12)) * f1 +
         (1 - 3 * del12 * del02 + del12 * (del02 * del02) + del02 * (del12 * del12)) * f2;
    dos = (-2 * del02 * ddel02 * del12 - ddel12 * (del02 * del02)) * f0 +
          (-ddel02 * (del12 * del12) - 2 * del02 * del12 * ddel12) * f1 +
          (-3 * del12 * ddel02 - 3 * del02 * ddel12 + 2 * del02 * del12 * ddel02 +
           ddel12 * (del02 * del02) + 2 * del02 * del12 * ddel12 + ddel02 * (del12 * del12)) * f2;
  }
  if (t[1].e == Ef) {
    const double e0 = t[0].e, f0 = t[0].f, e1 = t[2].e, f1 = t[2].f, e2 = t[3].e, f2 = t[3].f;
    const double del10 = (Ef - e0) / (e1 - e0);
    const double del20 = (Ef - e0) / (e2 - e0);
    const double ddel10 = 1 / (e1 - e0);
    const double ddel20 = 1 / (e2 - e0);
    const double ndp = (3 - del10 - del20) * f0 + del10 * f1 + del20 * f2;
    dos = ddel10 * del20 * ndp + del10 * ddel20 * ndp +
          del10 * del20 * ((-ddel10 - ddel20) * f0 + ddel10 * f1 + ddel20 * f2);
    nd = ndp * del10 * del20;
  }
  if (t[0].e == Ef) {
    dos = 0.0;
    nd = 0.0;
  }
}

// first index with E[i] >= x (upper = false) or E[i] > x (upper = true)
__device__ __forceinline__ int64_t bound(const double* E, int64_t nE, double x, bool upper) {
  int64_t lo = 0, hi = nE;
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if (upper ? (E[mid] <= x) : (E[mid] < x)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(256) k_lt(const double* freq, const double* func, int64_t nb,
                                            const int64_t* tri, int64_t ntri, const double* E,
                                            int64_t nE, double* dos, double* nd, double* step) {
#pragma clang fp contract(off)
  const int64_t total = ntri * nb;
  for (int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; id < total;
       id += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = id / nb, b = id % nb;
    Vtx v[3];
    double lo = 0, hi = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t k = tri[3 * t + c];
      v[c] = Vtx{freq[k * nb + b], func ? func[k * nb + b] : 1.0};
      lo = (c == 0) ? v[c].e : fmin(lo, v[c].e);
      hi = (c == 0) ? v[c].e : fmax(hi, v[c].e);
    }
    // energies inside [lo, hi]: the exact sort and closed forms
    const int64_t i0 = bound(E, nE, lo, false), i1 = bound(E, nE, hi, true);
    for (int64_t i = i0; i < i1; ++i) {
      const double Ef = E[i];
      Vtx a[4] = {Vtx{Ef, 0.0}, v[0], v[1], v[2]};
      sort4(a);
      double d, n;
      integrate_triangle(a, Ef, d, n);
      atomicAdd(dos + i, d);
      atomicAdd(nd + i, n);
    }
    // energies above: the sum of the three function values in the order the
    // sort leaves them when E is the largest
    if (i1 < nE) {
      Vtx a[4] = {Vtx{E[i1], 0.0}, v[0], v[1], v[2]};
      sort4(a);
      atomicAdd(step + i1, (a[0].f + a[1].f) + a[2].f);
    }
  }
}

// nd[i] += sum_{k <= i} step[k]; one workgroup, chunks of 1024
__global__ void __launch_bounds__(1024) k_step_scan(const double* step, double* nd, int64_t nE) {
  __shared__ double buf[1024];
  __shared__ double carry;
  if (threadIdx.x == 0) carry = 0.0;
  __syncthreads();
  for (int64_t base = 0; base < nE; base += 1024) {
    const int64_t i = base + threadIdx.x;
    buf[threadIdx.x] = (i < nE) ? step[i] : 0.0;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const double add = (threadIdx.x >= (unsigned)o) ? buf[threadIdx.x - o] : 0.0;
      __syncthreads();
      buf[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < nE) nd[i] += carry + buf[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
}

constexpr int kSmearTile = 1024;
constexpr double kPi32 = 3.1415927410125732;   // float32 pi, as the reference's literal

// grid: x = energy chunk of 256, y = q chunk of qper q-points
__global__ void __launch_bounds__(256) k_smear(int method, const double* freq, const double* w,
                                               int64_t nq, int64_t nb, double width, int64_t natom,
                                               const double* E, int64_t nE, int64_t qper,
                                               double* dos) {
#pragma clang fp contract(off)
  __shared__ double sf[kSmearTile];
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  const double Ei = (i < nE) ? E[i] : 0.0;
  const double Elo = E[blockIdx.x * (int64_t)256];
  const double Ehi = E[min(nE - 1, blockIdx.x * (int64_t)256 + 255)];
  const int64_t q0 = blockIdx.y * qper, q1 = min(nq, q0 + qper);
  const double cut = 40.0 * width;
  const double gnorm = sqrt(2 * kPi32) * width, hw = width / 2;
  double acc = 0.0;
  for (int64_t q = q0; q < q1; ++q) {
    double sq = 0.0;
    for (int64_t b0 = 0; b0 < nb; b0 += kSmearTile) {
      const int cnt = (int)min((int64_t)kSmearTile, nb - b0);
      const double* f = freq + q * nb + b0;
      // the q-point's frequencies are ascending: whole tiles out of reach of
      // a Gaussian are exact zeros
      if (method == 1 && (f[cnt - 1] < Elo - cut || f[0] > Ehi + cut)) continue;
      __syncthreads();
      for (int k = threadIdx.x; k < cnt; k += 256) sf[k] = f[k];
      __syncthreads();
      if (method == 1) {
        for (int k = 0; k < cnt; ++k) {
          const double x = (Ei - sf[k]) / width;
          sq += exp(-(x * x) / 2) / gnorm;
        }
      } else {
        for (int k = 0; k < cnt; ++k) {
          const double x = Ei - sf[k];
          sq += (hw / (x * x + hw * hw)) * (1 / kPi32);
        }
      }
    }
    acc += sq * w[q] / natom;
  }
  if (i < nE && acc != 0.0) atomicAdd(dos + i, acc);
}

void launch_dos_lt(const double* freq, const double* func, int64_t nb, const int64_t* tri,
                   int64_t ntri, const double* E, int64_t nE, double* dos, double* nd, double* step,
                   hipStream_t s) {
  launch_fill_zero(dos, nE * sizeof(double), s);
  launch_fill_zero(nd, nE * sizeof(double), s);
  launch_fill_zero(step, nE * sizeof(double), s);
  const int64_t total = ntri * nb;
  if (total > 0 && nE > 0) {
    const unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(total, 256), 65536);
    hipLaunchKernelGGL(k_lt, dim3(blocks), dim3(256), 0, s, freq, func, nb, tri, ntri, E, nE, dos,
                       nd, step);
    PPH_CHECK_LAUNCH();
  }
  if (nE > 0) {
    hipLaunchKernelGGL(k_step_scan, dim3(1), dim3(1024), 0, s, step, nd, nE);
    PPH_CHECK_LAUNCH();
  }
}

void launch_dos_smear(int method, const double* freq, const double* w, int64_t nq, int64_t nb,
                      double width, int64_t natom, const double* E, int64_t nE, double* dos,
                      hipStream_t s) {
  launch_fill_zero(dos, nE * sizeof(double), s);
  if (nq <= 0 || nb <= 0 || nE <= 0) return;
  const int64_t ex = ceil_div(nE, 256);
  // enough q chunks for ~16 workgroups per CU
  const int64_t qchunks = std::max<int64_t>(1, std::min<int64_t>(nq, ceil_div(4096, ex)));
  const int64_t qper = ceil_div(nq, qchunks);
  hipLaunchKernelGGL(k_smear, dim3((unsigned)ex, (unsigned)ceil_div(nq, qper)), dim3(256), 0, s,
                     method, freq, w, nq, nb, width, natom, E, nE, qper, dos);
  PPH_CHECK_LAUNCH();
}

}  // namespace pph
