// C ABI of libparphom_hip.so: handle life cycle and the batched per-q pipeline
//   K1 assembly -> tridiagonal reduction -> bisection / inverse iteration ->
//   back-transformation -> (FP32 path) FP64 Rayleigh-quotient refinement ->
//   velocity products -> unit conversion.
#include "../../include/parphom_hip.h"
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <array>
#include <type_traits>
#include <atomic>
#include <complex>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <exception>
#include <thread>
#include <vector>

namespace pph {

static thread_local std::string g_last_error;

// Automatic choice of the reduction (measured on MI355X, all eigenpairs):
// the two-stage path on the FP32 path for orders kTwoStageMinN ..
// kTwoStageMaxN, where its matrix-core band reduction outweighs the second
// back-transformation (n = 1092 x 256 q: 980 against 838 q-points/s; n = 2604:
// even; n = 4764, two batches of 32 in flight: 23.7 against 22.2).  Its panel
// factorisation and bulge chasing give each matrix one workgroup, so a single
// very large matrix stays on the one-stage path, and so does FP64 (its bulge
// chasing is register-bound: n = 4764, 16 q: 1.7 s against 1.1 s for the whole
// one-stage reduction).
constexpr int64_t kTwoStageMinN = 512, kTwoStageMaxN = 16384;

// Bumped by every device allocation and free of a DevBuf: a captured batch
// graph holds the addresses of its buffers and is replayed only while this
// count is the one it was captured at.
static std::atomic<uint64_t> g_alloc_gen{0};
// process default of Handle::cluster_rr (pph_debug_cluster_rr with no handle)
static std::atomic<int> g_cluster_rr_default{1};
// process default of Handle::vec_refine (pph_debug_vector_refinement with no handle)
static std::atomic<int> g_vec_refine_default{2};
// process default of Handle::vec_bf16
static std::atomic<int> g_vec_bf16_default{1};

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) {
      (void)hipFree(p);
      g_alloc_gen.fetch_add(1);
    }
    p = nullptr;
    bytes = 0;
  }
  void ensure(size_t b) {
    if (b <= bytes) return;
    release();
    if (b == 0) return;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      throw std::bad_alloc();
    }
    g_alloc_gen.fetch_add(1);
    bytes = b;
  }
  template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

struct Geometry {
  // host copies following lammps_read.F90:198-228 (linsolve / matinvtransp3)
  int64_t N = 0;
  double lat[9];                 // column-major, lat(i,k) at [i + 3k]
  std::vector<double> crys, pos; // N x 3 column-major
  double rec[9];                 // 2 pi inv(lat)^T, rows b_i, column-major
  std::vector<double> rmass;
};

static inline double L(const double* a, int i, int k) { return a[i + 3 * k]; }

// inverse of a column-major 3x3 by cofactors, ordered as lammps_read.F90:281-309
static void inv3(const double* A, double* B /* column-major */, double scale) {
  const double detinv =
      1.0 / (L(A, 0, 0) * L(A, 1, 1) * L(A, 2, 2) - L(A, 0, 0) * L(A, 1, 2) * L(A, 2, 1) -
             L(A, 0, 1) * L(A, 1, 0) * L(A, 2, 2) + L(A, 0, 1) * L(A, 1, 2) * L(A, 2, 0) +
             L(A, 0, 2) * L(A, 1, 0) * L(A, 2, 1) - L(A, 0, 2) * L(A, 1, 1) * L(A, 2, 0));
  auto S = [&](int i, int k, double v) { B[i + 3 * k] = v * scale; };
  S(0, 0, +detinv * (L(A, 1, 1) * L(A, 2, 2) - L(A, 1, 2) * L(A, 2, 1)));
  S(1, 0, -detinv * (L(A, 1, 0) * L(A, 2, 2) - L(A, 1, 2) * L(A, 2, 0)));
  S(2, 0, +detinv * (L(A, 1, 0) * L(A, 2, 1) - L(A, 1, 1) * L(A, 2, 0)));
  S(0, 1, -detinv * (L(A, 0, 1) * L(A, 2, 2) - L(A, 0, 2) * L(A, 2, 1)));
  S(1, 1, +detinv * (L(A, 0, 0) * L(A, 2, 2) - L(A, 0, 2) * L(A, 2, 0)));
  S(2, 1, -detinv * (L(A, 0, 0) * L(A, 2, 1) - L(A, 0, 1) * L(A, 2, 0)));
  S(0, 2, +detinv * (L(A, 0, 1) * L(A, 1, 2) - L(A, 0, 2) * L(A, 1, 1)));
  S(1, 2, -detinv * (L(A, 0, 0) * L(A, 1, 2) - L(A, 0, 2) * L(A, 1, 0)));
  S(2, 2, +detinv * (L(A, 0, 0) * L(A, 1, 1) - L(A, 0, 1) * L(A, 1, 0)));
}

// u = inv(A)^T x  (linsolve)
static void linsolve(const double* x, const double* A, double* u) {
  double B[9];
  inv3(A, B, 1.0);
  for (int k = 0; k < 3; ++k) u[k] = L(B, 0, k) * x[0] + L(B, 1, k) * x[1] + L(B, 2, k) * x[2];
}

// phase timers recorded on the handle's stream (HIP events)
// PH_HEMV: the k_hemv_sym launches alone (ms summed, count = launches)
// PH_HE2HB, PH_HB2ST, PH_Q2: the parts of the two-stage reduction and of its
// back-transformation (inside PH_TRID and PH_BACK)
enum Phase { PH_ASM = 0, PH_TRID, PH_TRIEIG, PH_BACK, PH_REFINE, PH_VEL, PH_TOTAL, PH_HEMV,
             PH_HE2HB, PH_HB2ST, PH_Q2, PH_N };

struct Timers {
  bool on = false;
  hipEvent_t ev[2 * PH_N] = {};
  bool used[PH_N] = {};
  double ms[PH_N] = {};
  int64_t cnt[PH_N] = {};
  void init() {
    if (ev[0]) return;
    for (auto& e : ev) PPH_HIP(hipEventCreate(&e));
  }
  void begin(int p, hipStream_t s) {
    if (!on) return;
    PPH_HIP(hipEventRecord(ev[2 * p], s));
    used[p] = true;
  }
  void end(int p, hipStream_t s) {
    if (!on) return;
    PPH_HIP(hipEventRecord(ev[2 * p + 1], s));
  }
  // per-launch event pairs for one kernel (the hemv of the reduction)
  std::vector<hipEvent_t> kev;
  int64_t kev_pending = 0;
  hipEvent_t* kernel_events(int64_t pairs) {
    if (!on) return nullptr;
    while ((int64_t)kev.size() < 2 * pairs) {
      hipEvent_t e;
      PPH_HIP(hipEventCreate(&e));
      kev.push_back(e);
    }
    kev_pending = pairs;
    return kev.data();
  }
  void collect() {
    if (!on) return;
    for (int64_t i = 0; i < kev_pending; ++i) {
      PPH_HIP(hipEventSynchronize(kev[2 * i + 1]));
      float t = 0.f;
      PPH_HIP(hipEventElapsedTime(&t, kev[2 * i], kev[2 * i + 1]));
      ms[PH_HEMV] += t;
      cnt[PH_HEMV] += 1;
    }
    kev_pending = 0;
    for (int p = 0; p < PH_N; ++p) {
      if (!used[p]) continue;
      PPH_HIP(hipEventSynchronize(ev[2 * p + 1]));
      float t = 0.f;
      PPH_HIP(hipEventElapsedTime(&t, ev[2 * p], ev[2 * p + 1]));
      ms[p] += t;
      cnt[p] += 1;
      used[p] = false;
    }
  }
  ~Timers() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : kev) (void)hipEventDestroy(e);
  }
};

struct BatchOut;

// The working state of one batch in flight: its stream, phase timers and
// device workspaces.  pph_run keeps several lanes busy at once (one host
// thread each), so that one batch's memory-bound reduction overlaps
// another's matrix-core products.
struct Lane {
  hipStream_t stream = nullptr;
  Timers tm;
  DevBuf qdev, A64, A32, dD, V, W, tau, dgl, egl, ypart, dpart, scal, part;
  DevBuf tdd, tee, te2, tilo, tihi, tinfo, tw, tzt, tfac, tinf;
  DevBuf Z, Z64, Tm, VT, Wt, vacc, num, den, lam, om, WT, C12, CG;
  DevBuf bv, Zt;           // block phase sums of D / dD, mode-contiguous vectors (block-sparse forms)
  DevBuf cmeta, cvec, cform;  // cluster Rayleigh-Ritz: cluster lists, combination vectors, their forms
  DevBuf vxt, vtol, vband; // vector correction: FP64 mode-contiguous vectors of a chunk, pair
                           // tolerances, the band of E (split first step)
  DevBuf gws, sws;         // partials of the GEMM dot epilogue / split-K and of the sparse forms
  DevBuf meta;             // per-matrix offsets and counts of the batch's outputs
  // two-stage reduction: band, bulge-chasing reflectors and their block
  // factors, panel T / Mx, Gram partials, stall flags
  DevBuf tsAB, tsV2, tsTau2, tsVTb, tsTp, tsG, tsInfo;
  // the batch captured as a HIP graph (range A): its shape key, the
  // allocation count it was captured at, and the host-side results of the
  // capture (counts, device pointers), which every replay reproduces
  struct Graph {
    hipGraphExec_t exec = nullptr;
    std::array<int64_t, 8> key{};
    uint64_t gen = 0;
  } graph;
  std::shared_ptr<BatchOut> graph_out;
  std::array<int64_t, 8> eager_key{};
  uint64_t eager_gen = ~uint64_t(0);
  bool graph_failed = false;
  int64_t graph_captures = 0, graph_replays = 0;
  Lane() = default;
  Lane(const Lane&) = delete;
  Lane& operator=(const Lane&) = delete;
  ~Lane() {
    if (graph.exec) (void)hipGraphExecDestroy(graph.exec);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// A handle is lane 0 plus the persistent per-system state.
struct Handle : Lane {
  int device = 0;
  pph_opts opts{};
  Geometry g;
  int64_t N = 0, n = 0;
  // leading dimension of the matrices the reduction works on (D64 / D32):
  // n rounded up to 16 elements, so that columns start on 128-byte lines
  int64_t ldA = 0;
  void set_n(int64_t nn) { n = nn; ldA = (nn + 15) / 16 * 16; }
  std::string err;
  // persistent device state
  DevBuf phi, crys, pos, lat, rmass, mask, cd_dc;
  double cd_norm = 0.0;
  // reduction: 0 one-stage (LATRD/HETRD), 1 two-stage (dense -> band ->
  // tridiagonal), -1 automatic; band: half-bandwidth of the two-stage band
  int reduction = -1;
  int band = -1;            // -1: the process default
  bool two_stage() const;
  // the one-launch reduction of small matrices (mode 2, or automatic) where
  // the order fits its LDS-resident kernel
  template <typename T> bool small_reduction() const;
  // the reduction a batch of this handle runs: 0 one-stage, 1 two-stage, 2 small
  int reduction_code() const;
  int band_width() const;
  bool graphs = true;      // replay captured batch graphs (range A)
  // FP32 path: Rayleigh-Ritz over clusters of close eigenvalues (range A)
  bool cluster_rr = g_cluster_rr_default.load() != 0;
  // FP32 path with eigenvectors: FP64 residual correction steps of the vectors (range A)
  int vec_refine = g_vec_refine_default.load();
  // the first correction step's two full products on the BF16 matrix cores
  bool vec_bf16 = g_vec_bf16_default.load() != 0;
  int gemm_tile = 0;       // GEMM tile shape of the eigensolver (kGemmTile*; 0: process default)
  int nb_panel = 32;       // reflectors per reduction panel
  int nb_back = 128;       // reflectors per back-transformation block (<= kVpad)
  int nchunk = 64;
  // Nonzero 3x3 blocks (a, b), a >= b, of Phi_sym listed by row a (sp_rowptr,
  // with sp_row / sp_col per block).  The quadratic forms of the refinement
  // and the velocities use them instead of dense products when at most a
  // quarter of all blocks are nonzero (a cutoff model in a large cell).
  DevBuf sp_rowptr, sp_row, sp_col, sp_phit;    // + the real block Phi~ of each
  DevBuf sp_tptr, sp_trow, sp_tblk;             // per atom a: the blocks (c, a), c > a
  DevBuf sp_flags;                              // N x N bytes: block (a, b) of Phi_sym nonzero
  int64_t sp_nblk = 0;
  bool sp_ok = false;
  int sp_mode = 1;         // 0 dense products, 1 automatic, 2 block lists whenever built
  // FP32 path with all eigenpairs and the FP64 correction of the vectors:
  // its velocity forms must be FP64 forms of the corrected FP64 vectors, which
  // the block lists give and the dense FP32 products of their FP32 rounding
  // do not (1e-6 of the largest velocity off at the acoustic modes of a small
  // cell, where omega divides the form)
  bool fp64_forms() const {
    return opts.precision == 'S' && opts.range == 'A' && opts.jobz == 'V' && vec_refine > 0;
  }
  bool use_sparse() const {
    return sp_nblk > 0 && (sp_mode == 2 || (sp_mode == 1 && (sp_ok || fp64_forms())));
  }
  // lanes 1.. (created on demand by pph_run)
  std::vector<std::unique_ptr<Lane>> more;
  Lane& lane(size_t i) { return i == 0 ? static_cast<Lane&>(*this) : *more[i - 1]; }
  size_t lanes() const { return 1 + more.size(); }
};


// Reduction chosen for handles whose `reduction` is -1 (automatic); set by
// pph_debug_reduction(NULL, ...), read by every handle (also the temporary
// ones pph_heevx makes).
static int g_reduction_default = -1;
static int g_band_default = 32;

int Handle::band_width() const { return band > 0 ? band : g_band_default; }

bool Handle::two_stage() const {
  const int r = reduction >= 0 ? reduction : g_reduction_default;
  if (r >= 0) return r == 1;
  return opts.precision == 'S' && n >= kTwoStageMinN && n <= kTwoStageMaxN;
}

template <typename T> bool Handle::small_reduction() const {
  const int r = reduction >= 0 ? reduction : g_reduction_default;
  return (r == -1 || r == 2) && n <= kSmallMaxN<T>;
}

int Handle::reduction_code() const {
  if (two_stage()) return 1;
  const bool small = opts.precision == 'S' ? small_reduction<float>() : small_reduction<double>();
  return small ? 2 : 0;
}

// ---------------------------------------------------------------- small dense FP64 (host)
// Eigenvalues of the Hermitian pencil (H, G), G positive definite, k <= 64:
// Cholesky G = L L^H, A = L^-1 H L^-H, cyclic complex Jacobi on A.
// Returns eigenvalues ascending and the vectors U (k x k, column-major) with
// U^H G U = I.
static void small_gen_eig(int k, std::vector<std::complex<double>> H, std::vector<std::complex<double>> G,
                          std::vector<double>& w, std::vector<std::complex<double>>& U) {
  using cd = std::complex<double>;
  auto at = [k](std::vector<cd>& M, int i, int j) -> cd& { return M[i + (size_t)j * k]; };
  // Cholesky (lower) in place in G
  for (int j = 0; j < k; ++j) {
    double d = at(G, j, j).real();
    for (int l = 0; l < j; ++l) d -= std::norm(at(G, j, l));
    d = std::sqrt(std::max(d, 1e-300));
    at(G, j, j) = d;
    for (int i = j + 1; i < k; ++i) {
      cd s = at(G, i, j);
      for (int l = 0; l < j; ++l) s -= at(G, i, l) * std::conj(at(G, j, l));
      at(G, i, j) = s / d;
    }
    for (int i = 0; i < j; ++i) at(G, i, j) = 0.0;
  }
  // A = L^-1 H L^-H : solve L X = H, then A^H = L^-1 X^H
  std::vector<cd> X(H);
  for (int c = 0; c < k; ++c)
    for (int i = 0; i < k; ++i) {
      cd s = at(X, i, c);
      for (int l = 0; l < i; ++l) s -= at(G, i, l) * at(X, l, c);
      at(X, i, c) = s / at(G, i, i);
    }
  std::vector<cd> A((size_t)k * k);
  for (int r = 0; r < k; ++r)          // row r of X -> solve L a = conj(X(r,:))^T
    for (int i = 0; i < k; ++i) {
      cd s = std::conj(at(X, r, i));
      for (int l = 0; l < i; ++l) s -= at(G, i, l) * at(A, l, r);
      at(A, i, r) = s / at(G, i, i);
    }
  for (int i = 0; i < k; ++i) {          // Hermitise
    at(A, i, i) = at(A, i, i).real();
    for (int j = i + 1; j < k; ++j) {
      const cd m = 0.5 * (at(A, i, j) + std::conj(at(A, j, i)));
      at(A, i, j) = m;
      at(A, j, i) = std::conj(m);
    }
  }
  std::vector<cd> Q((size_t)k * k, 0.0);
  for (int i = 0; i < k; ++i) at(Q, i, i) = 1.0;
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) (i == j ? tot : off) += std::norm(at(A, i, j));
    if (off <= 1e-32 * std::max(tot, 1e-300)) break;
    for (int p = 0; p < k; ++p)
      for (int q = p + 1; q < k; ++q) {
        const cd apq = at(A, p, q);
        const double ab = std::abs(apq);
        if (ab == 0.0) continue;
        const double app = at(A, p, p).real(), aqq = at(A, q, q).real();
        const cd em = std::conj(apq / ab);            // e^{-i phi}, a_pq = |a_pq| e^{i phi}
        const double tau = (aqq - app) / (2.0 * ab);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        // J = diag(1, e^{-i phi}) [c s; -s c] on the (p, q) plane; A <- J^H A J, Q <- Q J
        for (int r = 0; r < k; ++r) {
          const cd arp = at(A, r, p), arq = at(A, r, q);
          at(A, r, p) = c * arp - s * em * arq;
          at(A, r, q) = s * arp + c * em * arq;
        }
        for (int r = 0; r < k; ++r) {
          const cd apr = at(A, p, r), aqr = at(A, q, r);
          at(A, p, r) = c * apr - s * std::conj(em) * aqr;
          at(A, q, r) = s * apr + c * std::conj(em) * aqr;
        }
        for (int r = 0; r < k; ++r) {
          const cd qrp = at(Q, r, p), qrq = at(Q, r, q);
          at(Q, r, p) = c * qrp - s * em * qrq;
          at(Q, r, q) = s * qrp + c * em * qrq;
        }
      }
  }
  // U = L^-H Q
  U.assign((size_t)k * k, 0.0);
  for (int c = 0; c < k; ++c)
    for (int i = k - 1; i >= 0; --i) {
      cd s = at(Q, i, c);
      for (int l = i + 1; l < k; ++l) s -= std::conj(at(G, l, i)) * at(U, l, c);
      at(U, i, c) = s / at(G, i, i);
    }
  std::vector<int> ord(k);
  for (int i = 0; i < k; ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int x, int y) { return at(A, x, x).real() < at(A, y, y).real(); });
  w.resize(k);
  std::vector<cd> Us(U);
  for (int c = 0; c < k; ++c) {
    w[c] = at(A, ord[c], ord[c]).real();
    for (int i = 0; i < k; ++i) Us[i + (size_t)c * k] = at(U, i, ord[c]);
  }
  U.swap(Us);
}

// ---------------------------------------------------------------- eigen pipeline
// Householder vectors are stored explicitly with kVpad zero columns after the
// last one, so reflector blocks of any width up to kVpad read in bounds.
constexpr int64_t kVpad = 128;

struct EigResult {
  std::vector<int64_t> M;     // per matrix
  std::vector<int64_t> first; // per matrix: index of the first returned mode
  std::vector<int32_t> info;
  const int* info_dev = nullptr;  // device flags still to be or-ed into info (B ints), or null
  int64_t mcap = 0;           // leading dimension of w / columns of Z
  double* lam = nullptr;      // device, B x mcap (lambda)
  cplx<double>* Z64 = nullptr;  // device, B x n x mcap (when vectors)
  void* Zt = nullptr;           // device, vectors in the working precision
};

static int64_t range_mcap(const pph_opts& o, int64_t n) {
  if (o.range == 'I') return std::max<int64_t>(1, std::min(n, o.iu) - std::max<int64_t>(1, o.il) + 1);
  return n;
}

// Low-frequency modes after the Rayleigh quotients: their frequencies
// omega = K sqrt(lambda) magnify eigenvalue errors, so the lowest modes
// (omega < 50 cm^-1, at most 16 per matrix) get one Rayleigh-Ritz step over
// the span of all of them and a second-order correction
// -sum_j |x_j^H r|^2 / ((rho_j - theta) |x_j|^2) from every other computed
// mode j, all in FP64: W = D X_T, C1 = X^H W, C2 = X^H X_T on the GPU, the
// k x k algebra on the host.
static void complement_correction(Handle& h, Lane& L, int64_t B, const cplx<double>* D64,
                                  cplx<double>* X, int64_t mcap, int64_t K, int64_t mmax,
                                  const std::vector<int64_t>& Mv,
                                  const std::vector<std::complex<double>>& Uh,
                                  const std::vector<double>& thh, const std::vector<double>& invnb,
                                  std::vector<double>& corr, hipStream_t s);

static void refine_low_modes(Handle& h, Lane& L, int64_t B, const cplx<double>* D64, cplx<double>* X,
                             double* lam_dev, const double* nrm2_dev, const std::vector<int64_t>& Mv,
                             const std::vector<int64_t>& ilo, int64_t mcap, hipStream_t s) {
  const int64_t n = h.n;
  const double K32 = (double)(float)(15.633302f * 33.35641f);
  const double lam_low = (50.0 / K32) * (50.0 / K32);
  const int64_t KMAX = 16;
  std::vector<double> lam((size_t)B * mcap), nrm((size_t)B * mcap);
  PPH_HIP(hipMemcpyAsync(lam.data(), lam_dev, lam.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  PPH_HIP(hipMemcpyAsync(nrm.data(), nrm2_dev, nrm.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  PPH_HIP(hipStreamSynchronize(s));
  std::vector<int64_t> kb(B, 0);
  int64_t K = 0;
  for (int64_t b = 0; b < B; ++b) {
    int64_t k = 0;
    while (k < std::min(KMAX, Mv[b]) && lam[b * mcap + k] < lam_low) ++k;
    kb[b] = k;
    K = std::max(K, k);
  }
  if (K == 0) return;
  const int64_t Kp = 16;
  L.WT.ensure((size_t)B * n * Kp * sizeof(cplx<double>));
  L.C12.ensure((size_t)B * 2 * mcap * Kp * sizeof(cplx<double>));
  cplx<double>* WT = L.WT.as<cplx<double>>();
  cplx<double>* C1 = L.C12.as<cplx<double>>();
  cplx<double>* C2 = C1 + (size_t)B * mcap * Kp;
  int64_t mmax = 0;
  for (auto m : Mv) mmax = std::max(mmax, m);
  if (h.use_sparse()) {                      // W_T = D X(:, 0:K) from the listed blocks
    launch_sparse_spmm(h.sp_rowptr.as<int64_t>(), h.sp_col.as<int32_t>(), h.sp_tptr.as<int64_t>(),
                       h.sp_trow.as<int32_t>(), h.sp_tblk.as<int64_t>(), h.sp_nblk,
                       h.sp_phit.as<double>(), L.bv.as<cplx<double>>(), X, n * mcap, WT, n * Kp,
                       h.N, K, B, s);
  } else {
    GemmParams<double> g{};                 // W_T = D X(:, 0:K)
    g.opA = 'N'; g.opB = 'N'; g.m = n; g.n = K; g.k = n;
    g.alpha = cplx<double>{1, 0}; g.beta = cplx<double>{0, 0};
    g.A = D64; g.lda = h.ldA; g.strideA = h.ldA * n;
    g.B = X; g.ldb = n; g.strideB = n * mcap;
    g.C = WT; g.ldc = n; g.strideC = n * Kp;
    g.batch = B;
    launch_gemm<double>(g, s);
  }
  for (int which = 0; which < 2; ++which) {   // C1 = X^H W_T, C2 = X^H X_T
    GemmParams<double> g{};
    g.opA = 'C'; g.opB = 'N'; g.m = mmax; g.n = K; g.k = n;
    g.alpha = cplx<double>{1, 0}; g.beta = cplx<double>{0, 0};
    g.A = X; g.lda = n; g.strideA = n * mcap;
    g.B = which == 0 ? WT : X; g.ldb = n; g.strideB = which == 0 ? n * Kp : n * mcap;
    g.C = which == 0 ? C1 : C2; g.ldc = mcap; g.strideC = mcap * Kp;
    g.batch = B;
    launch_gemm<double>(g, s);
  }
  using cd = std::complex<double>;
  std::vector<cd> c1((size_t)B * mcap * Kp), c2((size_t)B * mcap * Kp);
  PPH_HIP(hipMemcpyAsync(c1.data(), C1, c1.size() * sizeof(cd), hipMemcpyDeviceToHost, s));
  PPH_HIP(hipMemcpyAsync(c2.data(), C2, c2.size() * sizeof(cd), hipMemcpyDeviceToHost, s));
  PPH_HIP(hipStreamSynchronize(s));

  // Ritz pairs of the low block, kept for the complement correction below
  std::vector<cd> Uh((size_t)B * Kp * Kp, 0.0);
  std::vector<double> thh((size_t)B * Kp, 0.0);
  std::vector<std::vector<double>> newls(B);
  for (int64_t b = 0; b < B; ++b) {
    const int64_t kt = kb[b], M = Mv[b];
    if (kt == 0) continue;
    const cd* C1b = c1.data() + (size_t)b * mcap * Kp;
    const cd* C2b = c2.data() + (size_t)b * mcap * Kp;
    double* lb = lam.data() + b * mcap;
    const double* nb = nrm.data() + b * mcap;
    std::vector<double> newl(lb, lb + kt);
    // modes within 1e-7 max|lambda| of this matrix of the last target join its cluster
    double tnorm_proxy = 0.0;
    for (int64_t k = 0; k < M; ++k) tnorm_proxy = std::max(tnorm_proxy, std::fabs(lb[k]));
    const double ctol = 1e-7 * std::max(tnorm_proxy, 1e-300);
    // one Rayleigh-Ritz block over all kt target modes: near-degenerate low
    // modes (the acoustic triplet at Gamma) may be closer to each other than
    // the FP32 residual, where a per-mode second-order correction is not
    // valid; Ritz values from the larger span are never worse, and what is
    // left couples only to modes above the 50 cm^-1 threshold
    int64_t c0 = 0;
    while (c0 < kt) {
      const int64_t c1e = kt;
      // extend over the full cluster in the whole spectrum (members beyond the targets
      // are not refined, but they are excluded from the coupling sums)
      int64_t cend = c1e;
      while (cend < M && lb[cend] - lb[cend - 1] <= ctol) ++cend;
      const int kc = (int)(c1e - c0);
      std::vector<cd> Hm((size_t)kc * kc), Gm((size_t)kc * kc);
      for (int i = 0; i < kc; ++i)
        for (int j = 0; j < kc; ++j) {
          Hm[i + (size_t)j * kc] = C1b[(c0 + i) + (c0 + j) * mcap];
          Gm[i + (size_t)j * kc] = C2b[(c0 + i) + (c0 + j) * mcap];
        }
      std::vector<double> th;
      std::vector<cd> U;
      small_gen_eig(kc, Hm, Gm, th, U);
      for (int m = 0; m < kc; ++m) {
        double corr = 0.0;
        for (int64_t j = 0; j < M; ++j) {
          if (j >= c0 && j < cend) continue;
          cd sjm = 0.0;
          for (int c = 0; c < kc; ++c)
            sjm += (C1b[j + (c0 + c) * mcap] - th[m] * C2b[j + (c0 + c) * mcap]) * U[c + (size_t)m * kc];
          const double dl = lb[j] - th[m];
          if (std::fabs(dl) > 0.0) corr -= std::norm(sjm) / (dl * nb[j]);
        }
        newl[c0 + m] = th[m] + corr;
        thh[b * Kp + m] = th[m];
        for (int c = 0; c < kc; ++c) Uh[(size_t)b * Kp * Kp + c + (size_t)m * Kp] = U[c + (size_t)m * kc];
      }
      c0 = c1e;
    }
    newls[b].swap(newl);
  }
  // The sums above run over the computed modes only.  When those start at
  // the bottom of the spectrum but stop short of its top (a range I or V
  // selection), the modes never computed contribute
  // -r_perp^H (D - theta)^-1 r_perp, r_perp the part of the residual outside
  // the computed span, on which D - theta is positive definite.  At the
  // acoustic triplet at Gamma (omega = K sqrt(lambda) near 0) that share
  // matters at the 1e-3 cm^-1 level for large n.
  std::vector<double> invnb((size_t)B * mcap, 0.0);
  bool any = false;
  for (int64_t b = 0; b < B; ++b) {
    const bool part = kb[b] > 0 && ilo[b] == 0 && Mv[b] < n;
    if (!part) {
      for (int64_t i = 0; i < Kp * Kp; ++i) Uh[(size_t)b * Kp * Kp + i] = 0.0;
      for (int64_t m = 0; m < Kp; ++m) thh[b * Kp + m] = 0.0;
      continue;
    }
    any = true;
    for (int64_t j = 0; j < Mv[b]; ++j)
      invnb[b * mcap + j] = nrm[b * mcap + j] > 0.0 ? 1.0 / nrm[b * mcap + j] : 0.0;
  }
  if (any) {
    std::vector<double> corr;
    complement_correction(h, L, B, D64, X, mcap, K, mmax, Mv, Uh, thh, invnb, corr, s);
    for (int64_t b = 0; b < B; ++b)
      if (kb[b] > 0 && ilo[b] == 0 && Mv[b] < n)
        for (int64_t m = 0; m < kb[b]; ++m) newls[b][m] += corr[b * Kp + m];
  }
  // the refined values replace the first kb[b] of each matrix in the host copy,
  // which then goes back in one transfer
  for (int64_t b = 0; b < B; ++b) {
    if (kb[b] == 0) continue;
    std::sort(newls[b].begin(), newls[b].end());
    std::copy(newls[b].begin(), newls[b].end(), lam.begin() + b * mcap);
  }
  PPH_HIP(hipMemcpyAsync(lam_dev, lam.data(), lam.size() * sizeof(double), hipMemcpyHostToDevice, s));
  PPH_HIP(hipStreamSynchronize(s));
}

// refine_low_modes for a batch whose matrices return all n modes (range A),
// with the selection, the k x k algebra and the correction sums on the GPU
// (refine.hip): the products with the lowest 16 vectors run for every batch,
// matrices without modes below the threshold leave the kernel at once.  No
// complement share: with every mode computed the sums are complete.
static void refine_low_modes_dev(Handle& h, Lane& L, int64_t B, const cplx<double>* D64,
                                 cplx<double>* X, double* lam_dev, const double* nrm2_dev,
                                 int64_t mcap, hipStream_t s) {
  const int64_t n = h.n;
  const double K32 = (double)(float)(15.633302f * 33.35641f);
  const double lam_low = (50.0 / K32) * (50.0 / K32);
  const int64_t Kp = 16, K = std::min<int64_t>(Kp, n);
  L.WT.ensure((size_t)B * n * Kp * sizeof(cplx<double>));
  L.C12.ensure((size_t)B * 2 * mcap * Kp * sizeof(cplx<double>));
  cplx<double>* WT = L.WT.as<cplx<double>>();
  cplx<double>* C1 = L.C12.as<cplx<double>>();
  cplx<double>* C2 = C1 + (size_t)B * mcap * Kp;
  if (h.use_sparse()) {                      // W_T = D X(:, 0:K) from the listed blocks
    launch_sparse_spmm(h.sp_rowptr.as<int64_t>(), h.sp_col.as<int32_t>(), h.sp_tptr.as<int64_t>(),
                       h.sp_trow.as<int32_t>(), h.sp_tblk.as<int64_t>(), h.sp_nblk,
                       h.sp_phit.as<double>(), L.bv.as<cplx<double>>(), X, n * mcap, WT, n * Kp,
                       h.N, K, B, s);
  } else {
    GemmParams<double> g{};                 // W_T = D X(:, 0:K)
    g.opA = 'N'; g.opB = 'N'; g.m = n; g.n = K; g.k = n;
    g.alpha = cplx<double>{1, 0}; g.beta = cplx<double>{0, 0};
    g.A = D64; g.lda = h.ldA; g.strideA = h.ldA * n;
    g.B = X; g.ldb = n; g.strideB = n * mcap;
    g.C = WT; g.ldc = n; g.strideC = n * Kp;
    g.batch = B;
    launch_gemm<double>(g, s);
  }
  for (int which = 0; which < 2; ++which) {   // C1 = X^H W_T, C2 = X^H X_T
    GemmParams<double> g{};
    g.opA = 'C'; g.opB = 'N'; g.m = n; g.n = K; g.k = n;
    g.alpha = cplx<double>{1, 0}; g.beta = cplx<double>{0, 0};
    g.A = X; g.lda = n; g.strideA = n * mcap;
    g.B = which == 0 ? WT : X; g.ldb = n; g.strideB = which == 0 ? n * Kp : n * mcap;
    g.C = which == 0 ? C1 : C2; g.ldc = mcap; g.strideC = mcap * Kp;
    g.batch = B;
    launch_gemm<double>(g, s);
  }
  launch_low_rr(lam_dev, nrm2_dev, C1, C2, n, mcap, lam_low, B, s);
}

// The complement share of the second-order correction for the Ritz pairs
// (theta_m, y_m = X(:, 0:K) U(:, m)) of each matrix, by conjugate gradients
// in FP64: r = D y - theta y, r_perp = P r with P = I - X diag(1/|x_j|^2) X^H
// over the computed modes (applied twice), then kCgIters steps on
// P (D - theta) P z = r_perp from z = 0; corr = -Re(r_perp^H z).  Matrices
// whose U and theta are zero get corr = 0.  Per step: one n x n x 16 product
// with D and two with X.  Columns M..mmax-1 of X (no mode) are cleared first:
// the batched products read them.
constexpr int kCgIters = 48;

static void complement_correction(Handle& h, Lane& L, int64_t B, const cplx<double>* D64,
                                  cplx<double>* X, int64_t mcap, int64_t K, int64_t mmax,
                                  const std::vector<int64_t>& Mv,
                                  const std::vector<std::complex<double>>& Uh,
                                  const std::vector<double>& thh, const std::vector<double>& invnb,
                                  std::vector<double>& corr, hipStream_t s) {
  using cd = std::complex<double>;
  const int64_t n = h.n, Kp = 16;
  const int64_t sv = n * Kp;                      // stride of the n x Kp blocks
  L.CG.ensure((size_t)B * (6 * sv + mcap * Kp + Kp * Kp) * sizeof(cd) +
              (size_t)B * (4 * Kp + mcap) * sizeof(double));
  cplx<double>* base = L.CG.as<cplx<double>>();
  cplx<double>* Y = base;
  cplx<double>* R = Y + B * sv;        // r_perp (kept)
  cplx<double>* Z = R + B * sv;
  cplx<double>* Rs = Z + B * sv;       // CG residual
  cplx<double>* P = Rs + B * sv;
  cplx<double>* Q = P + B * sv;
  cplx<double>* T = Q + B * sv;        // mcap x Kp
  cplx<double>* Ud = T + B * mcap * Kp;
  double* thd = reinterpret_cast<double*>(Ud + B * Kp * Kp);
  double* rho = thd + B * Kp;          // two buffers of B x Kp
  double* pq = rho + 2 * B * Kp;
  double* wn = pq + B * Kp;            // 1 / |x_j|^2, B x mcap
  PPH_HIP(hipMemcpyAsync(Ud, Uh.data(), Uh.size() * sizeof(cd), hipMemcpyHostToDevice, s));
  PPH_HIP(hipMemcpyAsync(thd, thh.data(), thh.size() * sizeof(double), hipMemcpyHostToDevice, s));
  PPH_HIP(hipMemcpyAsync(wn, invnb.data(), invnb.size() * sizeof(double), hipMemcpyHostToDevice, s));
  for (int64_t b = 0; b < B; ++b)
    if (Mv[b] < mmax)
      launch_fill_zero(X + b * n * mcap + Mv[b] * n, (mmax - Mv[b]) * n * (int64_t)sizeof(cd), s);
  auto gemm = [&](char opA, int64_t m, int64_t k, const cplx<double>* A, int64_t lda, int64_t sA,
                  const cplx<double>* Bm, int64_t ldb, int64_t sB, cplx<double>* C, int64_t ldc,
                  int64_t sC, double alpha, double beta) {
    GemmParams<double> g{};
    g.opA = opA; g.opB = 'N'; g.m = m; g.n = Kp; g.k = k;
    g.alpha = cplx<double>{alpha, 0}; g.beta = cplx<double>{beta, 0};
    g.A = A; g.lda = lda; g.strideA = sA;
    g.B = Bm; g.ldb = ldb; g.strideB = sB;
    g.C = C; g.ldc = ldc; g.strideC = sC;
    g.batch = B;
    launch_gemm<double>(g, s);
  };
  // V <- V - X diag(1/|x|^2) X^H V, twice
  auto project = [&](cplx<double>* V) {
    for (int pass = 0; pass < 2; ++pass) {
      gemm('C', mmax, n, X, n, n * mcap, V, n, sv, T, mcap, mcap * Kp, 1.0, 0.0);
      launch_scale_rows(T, wn, mmax, Kp, mcap, mcap * Kp, mcap, B, s);
      gemm('N', n, mmax, X, n, n * mcap, T, mcap, mcap * Kp, V, n, sv, -1.0, 1.0);
    }
  };
  // y = X(:, 0:K) U, r = (D X(:, 0:K)) U - theta y with D X(:, 0:K) in L.WT
  gemm('N', n, K, X, n, n * mcap, Ud, Kp, Kp * Kp, Y, n, sv, 1.0, 0.0);
  gemm('N', n, K, L.WT.as<cplx<double>>(), n, n * Kp, Ud, Kp, Kp * Kp, R, n, sv, 1.0, 0.0);
  launch_axpy_cols(R, Y, thd, -1.0, n, Kp, n, sv, Kp, B, s);
  project(R);
  const size_t vbytes = (size_t)B * sv * sizeof(cd);
  launch_fill_zero(Z, (int64_t)vbytes, s);
  PPH_HIP(hipMemcpyAsync(Rs, R, vbytes, hipMemcpyDeviceToDevice, s));
  PPH_HIP(hipMemcpyAsync(P, R, vbytes, hipMemcpyDeviceToDevice, s));
  launch_coldot_re(Rs, Rs, n, Kp, n, sv, B, rho, Kp, s);
  for (int it = 0; it < kCgIters; ++it) {
    double* rho_old = rho + (it % 2) * B * Kp;
    double* rho_new = rho + ((it + 1) % 2) * B * Kp;
    // q = P (D - theta) p
    if (h.use_sparse())
      launch_sparse_spmm(h.sp_rowptr.as<int64_t>(), h.sp_col.as<int32_t>(), h.sp_tptr.as<int64_t>(),
                         h.sp_trow.as<int32_t>(), h.sp_tblk.as<int64_t>(), h.sp_nblk,
                         h.sp_phit.as<double>(), L.bv.as<cplx<double>>(), P, sv, Q, sv, h.N, Kp,
                         B, s);
    else
      gemm('N', n, n, D64, h.ldA, h.ldA * n, P, n, sv, Q, n, sv, 1.0, 0.0);
    launch_axpy_cols(Q, P, thd, -1.0, n, Kp, n, sv, Kp, B, s);
    project(Q);
    launch_coldot_re(P, Q, n, Kp, n, sv, B, pq, Kp, s);
    launch_cg_update(Z, Rs, P, Q, rho_old, pq, n, Kp, n, sv, Kp, B, s);
    launch_coldot_re(Rs, Rs, n, Kp, n, sv, B, rho_new, Kp, s);
    launch_cg_direction(P, Rs, rho_new, rho_old, n, Kp, n, sv, Kp, B, s);
  }
  launch_coldot_re(R, Z, n, Kp, n, sv, B, pq, Kp, s);
  corr.assign((size_t)B * Kp, 0.0);
  PPH_HIP(hipMemcpyAsync(corr.data(), pq, corr.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  PPH_HIP(hipStreamSynchronize(s));
  for (auto& c : corr) c = -c;
}

// FP64 residual correction of the FP32 path's vectors (vecref.hip), range A:
// h.vec_refine steps of X <- X + X E with E from M = X^H (D X - X Theta)
// (the first with full products, later ones over a band of close modes),
// on the FP64 vectors Z64 and their FP32 copy L.Z (both updated), values
// theta (the refined quotients, kept) and |x|^2 in L.den (refreshed);
// fp32_exact: X64 still equals the FP32 vectors of L.Z and L.Zt.  R32
// and Em are scratch of B n mcap complex64 each, the caller's dead buffers
// (the reduced matrices and the tridiagonal vectors).  The residuals and the
// velocity forms run over chunks of kVecRefChunk matrices (one FP64
// mode-contiguous copy of a chunk); the products over the whole batch.  With
// velocities and block-sparse forms, vacc gets the forms of the corrected
// vectors.
constexpr int64_t kVecRefChunk = 8;

// Rayleigh-Ritz over the clusters of L.cmeta (refine.hip): the forms of the
// pairs' combination vectors (block-sparse, or the GEMM's Hermitian-dot
// epilogue with D64) give the off-diagonal entries of X^H D X and, with v,
// of X^H dD X; the diagonal ones are L.num and L.vacc.  The vectors come from
// src (element (r, k) of matrix b at b n mcap + r srs + k sks): the FP32 ones
// or the corrected FP64 ones.  Writes the Ritz values to theta, the Ritz
// vectors to Z64 and L.Z, the Ritz velocity forms to L.vacc (with v).
// The combination vectors go to `scratch` (a dead buffer of scratch_bytes)
// when it is large enough, else to L.cvec.
template <typename Ts>
static void cluster_rayleigh_ritz(Handle& h, Lane& L, int64_t B, const cplx<double>* D64,
                                  const cplx<Ts>* src, int64_t srs, int64_t sks, bool v,
                                  double* theta, int64_t mcap, void* scratch,
                                  size_t scratch_bytes, hipStream_t s) {
  const int64_t n = h.n;
  const bool forms = h.use_sparse();
  constexpr int64_t P = kClusterSlots;      // per-matrix stride of the slot forms
  const int64_t S = cluster_slots(n);          // slots in use
  const size_t cv_bytes = (size_t)B * n * S * sizeof(cplx<double>);
  if (scratch_bytes < cv_bytes) L.cvec.ensure(cv_bytes);
  L.cform.ensure((size_t)B * 4 * P * sizeof(double));
  const int32_t* meta = L.cmeta.as<int32_t>();
  cplx<double>* cv = scratch_bytes >= cv_bytes ? static_cast<cplx<double>*>(scratch)
                                               : L.cvec.as<cplx<double>>();
  double* numc = L.cform.as<double>();
  double* vc = numc + B * P;
  v = v && forms;
  launch_fill_zero(numc, B * 4 * P * (int64_t)sizeof(double), s);
  if (forms) {
    launch_cluster_gather<Ts>(meta, src, srs, sks, n * mcap, cv, S, 1, n, S, B, s);
    L.sws.ensure((size_t)sparse_dots_ws_bytes(h.N, S, B, v ? 4 : 1));
    launch_sparse_dots<double>(h.sp_rowptr.as<int64_t>(), h.sp_col.as<int32_t>(), h.sp_nblk,
                               h.sp_phit.as<double>(), L.bv.as<cplx<double>>(), 0, v ? 4 : 1, cv, S,
                               n * S, h.N, S, B, numc, P, v ? vc : nullptr, P, 3 * P,
                               L.sws.as<double>(), s);
  } else {
    launch_cluster_gather<Ts>(meta, src, srs, sks, n * mcap, cv, 1, n, n, S, B, s);
    GemmParams<double> g{};
    g.opA = 'N'; g.opB = 'N';
    g.m = n; g.n = S; g.k = n;
    g.alpha = cplx<double>{1, 0};
    g.A = D64; g.lda = h.ldA; g.strideA = h.ldA * n;
    g.B = cv; g.ldb = n; g.strideB = n * S;
    g.batch = B;
    g.dot_out = numc; g.dot_stride = P;
    g.herm_dot = 1;
    g.dotX = cv; g.lddx = n; g.strideX = n * S;
    L.gws.ensure((size_t)gemm_ws_bytes(g));
    g.ws = L.gws.p;
    launch_gemm<double>(g, s);
  }
  launch_cluster_rr(meta, L.Z64.as<cplx<double>>(), L.Z.as<cplx<float>>(), n, mcap, theta,
                    L.den.as<double>(), L.num.as<double>(), numc, v ? L.vacc.as<double>() : nullptr,
                    vc, B, s);
}

static void correct_vectors(Handle& h, Lane& L, int64_t B, const cplx<double>* D64,
                            cplx<float>* R32, cplx<float>* Em, double* theta, int64_t mmax,
                            int64_t mcap, bool want_num, bool fp32_exact, hipStream_t s) {
  const int64_t n = h.n;
  const bool forms = h.use_sparse();
  if (!forms && !D64) throw std::runtime_error("correct_vectors: no FP64 matrices");
  // the FP64 copy of a chunk of matrices (mode-contiguous vectors, or D X
  // for dense D): in L.Zt (the FP32 mode-contiguous vectors, read only
  // before this step; room for half the batch) when it holds the chunk,
  // else in a buffer of its own.  A chunk is as many matrices as L.Zt or
  // 1 GiB holds, at least kVecRefChunk, so small matrices run in one or two
  // chunks instead of many launches of a few matrices each.
  const size_t per_mat = (size_t)n * mcap * sizeof(cplx<double>);
  const size_t room = std::max<size_t>(L.Zt.bytes, (size_t)1 << 30);
  const int64_t Bc = std::min<int64_t>(B, std::max<int64_t>(kVecRefChunk, (int64_t)(room / per_mat)));
  const size_t xt_bytes = (size_t)Bc * per_mat;
  if (L.Zt.bytes < xt_bytes) L.vxt.ensure(xt_bytes);
  L.vtol.ensure((size_t)B * sizeof(double));
  cplx<double>* X64 = L.Z64.as<cplx<double>>();
  cplx<float>* X32 = L.Z.as<cplx<float>>();
  cplx<double>* Xt = L.Zt.bytes >= xt_bytes ? L.Zt.as<cplx<double>>() : L.vxt.as<cplx<double>>();
  double* den = L.den.as<double>();
  const int64_t sz = n * mcap;              // per-matrix stride of the n x mcap arrays
  launch_lam_tol(theta, mmax, mcap, kClusterTol, L.vtol.as<double>(), B, s);
  for (int step = 0; step < h.vec_refine; ++step) {
    // R = D X - X Theta in FP64, rounded to FP32.  Before the first step the
    // vectors are still the FP32 ones, whose mode-contiguous copy L.Zt holds
    // (the block-sparse Rayleigh numerators read it): the whole batch at once
    if (step == 0 && fp32_exact && forms) {
      launch_sparse_residual<float>(h.sp_rowptr.as<int64_t>(), h.sp_col.as<int32_t>(),
                                    h.sp_tptr.as<int64_t>(), h.sp_trow.as<int32_t>(),
                                    h.sp_tblk.as<int64_t>(), h.sp_nblk, h.sp_phit.as<double>(),
                                    L.bv.as<cplx<double>>(), L.Zt.as<cplx<float>>(), mcap, sz, theta,
                                    mcap, h.N, mmax, R32, mcap, sz, B, s);
    } else for (int64_t b0 = 0; b0 < B; b0 += Bc) {
      const int64_t bc = std::min(Bc, B - b0);
      if (forms) {
        launch_transpose<double>(X64 + b0 * sz, n, mmax, n, sz, Xt, mcap, sz, bc, s);
        launch_sparse_residual<double>(h.sp_rowptr.as<int64_t>(), h.sp_col.as<int32_t>(),
                               h.sp_tptr.as<int64_t>(), h.sp_trow.as<int32_t>(),
                               h.sp_tblk.as<int64_t>(), h.sp_nblk, h.sp_phit.as<double>(),
                               L.bv.as<cplx<double>>() + b0 * h.sp_nblk * 4, Xt, mcap, sz,
                               theta + b0 * mcap, mcap, h.N, mmax, R32 + b0 * sz, mcap, sz, bc, s);
      } else {
        GemmParams<double> g{};
        g.opA = 'N'; g.opB = 'N';
        g.m = n; g.n = mmax; g.k = n;
        g.alpha = cplx<double>{1, 0}; g.beta = cplx<double>{0, 0};
        g.A = D64 + b0 * h.ldA * n; g.lda = h.ldA; g.strideA = h.ldA * n;
        g.B = X64 + b0 * sz; g.ldb = n; g.strideB = sz;
        g.C = Xt; g.ldc = n; g.strideC = sz;
        g.batch = bc;
        // an FP64 product: the 128 x 128 tile of the FP32 ones does not apply
        if (gemm_current_tile() == kGemmTile128x128) g.tile = kGemmTile64x64;
        launch_gemm<double>(g, s);
        launch_dense_residual(Xt, X64 + b0 * sz, theta + b0 * mcap, mcap, n, mmax, sz, R32 + b0 * sz,
                              bc, s);
      }
    }
    const int64_t smb = mcap * (2 * kVecRefBand + 1);
    const bool band_ok = mmax > 4 * kVecRefBand;
    // the band of kVecRefBand modes on either side of the diagonal in FP32:
    // Eb = the band of E in L.vband
    auto band_part = [&](cplx<float>* Eb) {
      if (forms)
        launch_band_gram(X32, n, sz, R32, mcap, 1, sz, true, n, mmax, Eb, smb, B, s);
      else
        launch_band_gram(X32, n, sz, R32, 1, n, sz, false, n, mmax, Eb, smb, B, s);
      launch_band_coeffs(Eb, smb, mmax, theta, den, mcap, L.vtol.as<double>(), B, s);
    };
    // after the first step only pairs of close modes carry a noticeable
    // error: the band alone
    if (step > 0 && band_ok) {
      band_part(Em);
      launch_band_apply(X32, n, sz, Em, smb, n, mmax, R32, sz, false, B, s);
      launch_vecref_update(X64, X32, R32, n * mmax, sz, B, s);
      launch_colnorm2<double>(X64, n, mmax, n, sz, B, den, mcap, s);
      continue;
    }
    // the first step: full products.  With BF16 products, the band (whose
    // pairs of close modes carry the large coefficients) comes from the FP32
    // banded kernels and the BF16 products supply only the far pairs, whose
    // coefficients need a few digits
    const bool split = h.vec_bf16 && band_ok;
    cplx<float>* Eband = nullptr;
    if (split) {
      L.vband.ensure((size_t)B * smb * sizeof(cplx<float>));
      Eband = L.vband.as<cplx<float>>();
      band_part(Eband);
    }
    // M = X^H R on the matrix cores (R stored as R^H, mode-contiguous, on
    // the block-sparse route)
    {
      GemmParams<float> g{};
      g.opA = 'C'; g.opB = forms ? 'C' : 'N';
      g.m = mmax; g.n = mmax; g.k = n;
      g.alpha = cplx<float>{1, 0}; g.beta = cplx<float>{0, 0};
      g.A = X32; g.lda = n; g.strideA = sz;
      g.B = R32; g.ldb = forms ? mcap : n; g.strideB = sz;
      g.C = Em; g.ldc = mcap; g.strideC = mcap * mcap;
      g.batch = B;
      if (split) launch_cgemm_bf16(g, s);
      else launch_gemm<float>(g, s);
    }
    launch_vecref_coeffs(Em, mcap, mcap * mcap, mmax, theta, den, mcap, L.vtol.as<double>(),
                         split ? kVecRefBand : -1, B, s);
    // W = X E (into R32), then X += W and the FP32 copy
    {
      GemmParams<float> g{};
      g.opA = 'N'; g.opB = 'N';
      g.m = n; g.n = mmax; g.k = mmax;
      g.alpha = cplx<float>{1, 0}; g.beta = cplx<float>{0, 0};
      g.A = X32; g.lda = n; g.strideA = sz;
      g.B = Em; g.ldb = mcap; g.strideB = mcap * mcap;
      g.C = R32; g.ldc = n; g.strideC = sz;
      g.batch = B;
      if (split) launch_cgemm_bf16(g, s);
      else launch_gemm<float>(g, s);
    }
    if (split) launch_band_apply(X32, n, sz, Eband, smb, n, mmax, R32, sz, true, B, s);
    launch_vecref_update(X64, X32, R32, n * mmax, sz, B, s);
    launch_colnorm2<double>(X64, n, mmax, n, sz, B, den, mcap, s);
  }
  // the forms of the corrected vectors: with velocities and block-sparse
  // forms, the velocity forms; with want_num, the Rayleigh numerators (the
  // diagonal of X^H D X for a cluster step that follows).  In FP64 on the
  // FP64 vectors: the cluster step combines them with forms of FP64
  // combinations of the same vectors, whose differences must not carry the
  // FP32 rounding of either
  const bool v = forms && h.opts.vel != 0;
  if (want_num) launch_fill_zero(L.num.p, B * mcap * (int64_t)sizeof(double), s);
  if (forms && (v || want_num)) {
    // timed as the velocity phase (inside the refinement phase)
    if (v) L.tm.begin(PH_VEL, s);
    const int op0 = want_num ? 0 : 1, nops = (want_num ? 1 : 0) + (v ? 3 : 0);
    if (v) {
      L.vacc.ensure((size_t)B * mcap * 6 * sizeof(double));
      launch_fill_zero(L.vacc.p, B * mcap * 3 * (int64_t)sizeof(double), s);
    }
    L.sws.ensure((size_t)sparse_dots_ws_bytes(h.N, mmax, Bc, nops));
    for (int64_t b0 = 0; b0 < B; b0 += Bc) {
      const int64_t bc = std::min(Bc, B - b0);
      launch_transpose<double>(X64 + b0 * sz, n, mmax, n, sz, Xt, mcap, sz, bc, s);
      launch_sparse_dots<double>(h.sp_rowptr.as<int64_t>(), h.sp_col.as<int32_t>(), h.sp_nblk,
                                 h.sp_phit.as<double>(), L.bv.as<cplx<double>>() + b0 * h.sp_nblk * 4,
                                 op0, nops, Xt, mcap, sz, h.N, mmax, bc,
                                 want_num ? L.num.as<double>() + b0 * mcap : nullptr, mcap,
                                 v ? L.vacc.as<double>() + b0 * 3 * mcap : nullptr, mcap, 3 * mcap,
                                 L.sws.as<double>(), s);
    }
    if (v) L.tm.end(PH_VEL, s);
  } else if (want_num) {
    GemmParams<double> g{};
    g.opA = 'N'; g.opB = 'N';
    g.m = n; g.n = mmax; g.k = n;
    g.alpha = cplx<double>{1, 0};
    g.A = D64; g.lda = h.ldA; g.strideA = h.ldA * n;
    g.B = X64; g.ldb = n; g.strideB = sz;
    g.batch = B;
    g.dot_out = L.num.as<double>(); g.dot_stride = mcap;
    g.herm_dot = 1;
    g.dotX = X64; g.lddx = n; g.strideX = sz;
    L.gws.ensure((size_t)gemm_ws_bytes(g));
    g.ws = L.gws.p;
    launch_gemm<double>(g, s);
  }
}

// A: B matrices (n x n, working precision, destroyed); D64: FP64 matrices for
// the refinement (FP32 path only).  want_vec: caller needs eigenvectors.
template <typename T>
static void eig_batch(Handle& h, Lane& L, int64_t B, cplx<T>* A, const cplx<double>* D64, bool want_vec,
                      double vl_l, double vu_l, EigResult& R) {
  const int64_t n = h.n;
  hipStream_t s = L.stream;
  const pph_opts& o = h.opts;
  const bool fp32 = std::is_same<T, float>::value;
  const bool vectors = want_vec || fp32;      // the FP32 path refines with its vectors
  const gemm_tile_scope tile_scope(h.gemm_tile);
  const int nb = h.nb_panel;
  const int64_t mcap = range_mcap(o, n);
  R.mcap = mcap;

  // ---- reduction to a real tridiagonal
  const bool two = h.two_stage();
  const int64_t bwd = h.band_width();
  const int64_t kmax2 = (n > 1) ? hb2st_tasks(n, bwd, 0) : 1;
  const int64_t G2 = (n > 1) ? ceil_div(n - 1, bwd) : 1;
  L.V.ensure((size_t)B * n * (n + kVpad) * sizeof(cplx<T>));
  L.W.ensure((size_t)B * n * std::max<int64_t>(nb, 2 * bwd) * sizeof(cplx<T>));
  L.tau.ensure((size_t)B * n * sizeof(cplx<T>));
  L.dgl.ensure((size_t)B * n * sizeof(T));
  L.egl.ensure((size_t)B * n * sizeof(T));
  TridiagWork<T> tw{};
  tw.n = n; tw.lda = h.ldA; tw.batch = B; tw.nb = nb;
  tw.A = A; tw.strideA = h.ldA * n;
  tw.V = L.V.as<cplx<T>>(); tw.strideV = n * (n + kVpad);
  tw.W = L.W.as<cplx<T>>(); tw.strideW = n * nb;
  tw.tau = L.tau.as<cplx<T>>(); tw.strideTau = n;
  tw.d = L.dgl.as<T>(); tw.e = L.egl.as<T>(); tw.strideDE = n;
  TwoStageWork<T> ts{};
  const bool small = !two && h.small_reduction<T>();
  L.tm.begin(PH_TRID, s);
  if (small) {
    tridiagonalise_small<T>(tw, n + kVpad, s);
  } else if (two) {
    constexpr int64_t kChunks = 16;
    L.tsAB.ensure((size_t)B * n * 2 * bwd * sizeof(cplx<T>));
    L.tsV2.ensure((size_t)B * n * kmax2 * bwd * sizeof(cplx<T>));
    L.tsTau2.ensure((size_t)B * n * kmax2 * sizeof(cplx<T>));
    L.tsTp.ensure((size_t)B * 2 * bwd * bwd * sizeof(cplx<T>));
    L.tsG.ensure((size_t)B * kChunks * bwd * bwd * sizeof(cplx<T>));
    L.tsInfo.ensure((size_t)B * sizeof(int));
    launch_fill_zero(L.tsInfo.p, B * sizeof(int), s);
    ts.n = n; ts.lda = h.ldA; ts.batch = B; ts.bw = bwd;
    ts.A = A; ts.strideA = h.ldA * n;
    ts.V = tw.V; ts.strideV = tw.strideV;
    ts.tau = tw.tau; ts.strideTau = n;
    ts.X = L.W.as<cplx<T>>(); ts.VT = ts.X + n * bwd; ts.strideXV = 2 * n * bwd;
    ts.Tp = L.tsTp.as<cplx<T>>(); ts.strideTp = 2 * bwd * bwd;
    ts.gpart = L.tsG.as<cplx<T>>(); ts.nchunk = kChunks;
    ts.AB = L.tsAB.as<cplx<T>>(); ts.strideAB = n * 2 * bwd;
    ts.V2 = L.tsV2.as<cplx<T>>(); ts.tau2 = L.tsTau2.as<cplx<T>>(); ts.kmax = kmax2;
    ts.strideV2 = n * kmax2 * bwd; ts.strideTau2 = n * kmax2;
    ts.d = tw.d; ts.e = tw.e; ts.strideDE = n;
    ts.info = L.tsInfo.as<int>();
    L.tm.begin(PH_HE2HB, s);
    he2hb<T>(ts, s);
    L.tm.end(PH_HE2HB, s);
    L.tm.begin(PH_HB2ST, s);
    hb2st<T>(ts, s);
    L.tm.end(PH_HB2ST, s);
  } else {
    // the hemv's per-workgroup partials: 2 ntmax slots of ntmax * 64 rows
    const int64_t ntmax = hemv_tiles(n), ldY = ntmax * 64;
    L.ypart.ensure((size_t)B * 2 * ntmax * ldY * sizeof(cplx<T>));
    L.dpart.ensure((size_t)B * ntmax * 2 * nb * sizeof(cplx<T>));
    L.scal.ensure((size_t)B * 256 * sizeof(double));
    const int64_t nparts = n / 64 + 2;
    L.part.ensure((size_t)B * 3 * nparts * sizeof(double));
    tw.ypart = L.ypart.as<cplx<T>>(); tw.strideY = 2 * ntmax * ldY; tw.ldY = ldY; tw.ntmax = ntmax;
    tw.dpart = L.dpart.as<cplx<T>>(); tw.strideDp = ntmax * 2 * nb;
    tw.scal = L.scal.as<double>(); tw.strideScal = 256;
    tw.part = L.part.as<double>(); tw.strideP = 3 * nparts; tw.yvOff = nparts;
    tw.nchunk = h.nchunk;
    tw.hemv_ev = L.tm.kernel_events(ceil_div(n - 1, kHemvTimeStride));
    tridiagonalise<T>(tw, s);
  }
  L.tm.end(PH_TRID, s);

  // ---- tridiagonal eigenproblem in FP64
  L.tdd.ensure((size_t)B * n * sizeof(double));
  L.tee.ensure((size_t)B * n * sizeof(double));
  L.te2.ensure((size_t)B * n * sizeof(double));
  L.tilo.ensure((size_t)B * sizeof(int64_t));
  L.tihi.ensure((size_t)B * sizeof(int64_t));
  L.tinfo.ensure((size_t)B * 4 * sizeof(double));
  L.tw.ensure((size_t)B * mcap * sizeof(double));
  L.tinf.ensure((size_t)B * sizeof(int));
  TriEigWork te{};
  te.n = n; te.batch = B;
  te.d = tw.d; te.e = tw.e; te.strideDE = n; te.de_is_float = fp32 ? 1 : 0;
  te.dd = L.tdd.as<double>(); te.ee = L.tee.as<double>(); te.e2 = L.te2.as<double>();
  te.strideT = n;
  te.range = o.range; te.vl = vl_l; te.vu = vu_l; te.il = o.il; te.iu = o.iu;
  te.abstol = o.abstol; te.orfac = o.orfac;
  // the FP32 path's eigenvalues come from the FP64 refinement; bisection on
  // its FP32 tridiagonal only supplies the inverse-iteration shifts
  te.rtol = fp32 ? 1e-13 : 0.0;
  // with shifts within 1e-13 |T| and pairs closer than 1e-9 |T| handled as
  // clusters, two solves leave at most (1e-13 / 1e-9)^2 of another vector in
  // an FP32-path vector, below what its FP32 back-transformation keeps
  te.nsolve = fp32 ? 2 : 3;
  te.ilo = L.tilo.as<int64_t>(); te.ihi = L.tihi.as<int64_t>();
  te.tinfo = L.tinfo.as<double>();
  te.w = L.tw.as<double>(); te.ldw = mcap;
  te.info = L.tinf.as<int>();
  te.want_vectors = vectors ? 1 : 0;
  L.tm.begin(PH_TRIEIG, s);
  tri_prepare(te, s);
  // range A: every matrix returns its n modes from the bottom, so the counts
  // are known here and the batch runs without a host round trip (a stalled
  // bulge-chasing pipeline's flags stay on the device until the outputs are
  // written); ranges V and I read the Sturm-count ranges back
  const bool all_modes = o.range == 'A';
  std::vector<int64_t> ilo(B, 0), ihi(B, n);
  std::vector<int> rinfo(B, 0);
  if (!all_modes) {
    PPH_HIP(hipMemcpyAsync(ilo.data(), te.ilo, B * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PPH_HIP(hipMemcpyAsync(ihi.data(), te.ihi, B * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    if (two)
      PPH_HIP(hipMemcpyAsync(rinfo.data(), ts.info, B * sizeof(int), hipMemcpyDeviceToHost, s));
    PPH_HIP(hipStreamSynchronize(s));
  }
  R.M.assign(B, 0);
  R.first.assign(B, 0);
  // a stalled bulge-chasing pipeline reports as a convergence failure
  R.info.assign(rinfo.begin(), rinfo.end());
  R.info_dev = (all_modes && two) ? ts.info : nullptr;
  int64_t mmax = 0;
  for (int64_t b = 0; b < B; ++b) {
    R.M[b] = std::min(mcap, ihi[b] - ilo[b]);
    mmax = std::max(mmax, R.M[b]);
  }
  te.mmax = std::max<int64_t>(mmax, 1);
  tri_bisect(te, s);
  R.lam = te.w;
  if (!(vectors && mmax > 0)) L.tm.end(PH_TRIEIG, s);

  if (vectors && mmax > 0) {
    const int64_t ldzt = mcap;
    L.tzt.ensure((size_t)B * n * ldzt * sizeof(double));
    L.tfac.ensure((size_t)B * tri_fac_doubles(n, ldzt) * sizeof(double));
    te.zt = L.tzt.as<double>(); te.ldzt = ldzt; te.strideZt = n * ldzt;
    te.fac = L.tfac.as<double>(); te.strideFac = tri_fac_doubles(n, ldzt);
    tri_vectors(te, s);
    L.tm.end(PH_TRIEIG, s);

    // ---- back-transformation, in blocks of nbt reflectors
    const int64_t nbt = h.nb_back;
    // the Gram products read V columns up to npanel * nbt, nbt - 1 past the last
    if (nbt > kVpad) throw std::runtime_error("back-transformation block wider than the V padding");
    L.Z.ensure((size_t)B * n * mcap * sizeof(cplx<T>));
    const int64_t strideWs = (n + nbt) * nbt;
    L.Tm.ensure((size_t)B * strideWs * sizeof(cplx<T>));
    BackWork<T> bw{};
    bw.n = n; bw.batch = B; bw.nb = nbt; bw.m = mmax;
    bw.V = tw.V; bw.strideV = tw.strideV;
    bw.tau = tw.tau; bw.strideTau = n;
    bw.zt = te.zt; bw.ldzt = ldzt; bw.strideZt = te.strideZt;
    bw.Z = L.Z.as<cplx<T>>(); bw.ldz = n; bw.strideZ = n * mcap;
    bw.Tm = L.Tm.as<cplx<T>>(); bw.VT = L.VT.as<cplx<T>>(); bw.Wt = L.Wt.as<cplx<T>>();
    bw.strideWs = std::max<int64_t>(strideWs, std::max<int64_t>(n * nbt, nbt * mcap));
    // Tm/VT/Wt were sized with the same per-matrix stride
    L.Tm.ensure((size_t)B * bw.strideWs * sizeof(cplx<T>));
    L.VT.ensure((size_t)B * bw.strideWs * sizeof(cplx<T>));
    L.Wt.ensure((size_t)B * bw.strideWs * sizeof(cplx<T>));
    bw.Tm = L.Tm.as<cplx<T>>(); bw.VT = L.VT.as<cplx<T>>(); bw.Wt = L.Wt.as<cplx<T>>();
    L.gws.ensure((size_t)back_transform_ws_bytes<T>(n, B, nbt, mmax, two ? bwd : 1));
    bw.ws = L.gws.p;
    L.tm.begin(PH_BACK, s);
    if (two) {
      // Z = Q1 (Q2 Z_T): the bulge-chasing reflectors first, then the he2hb
      // panels (their reflectors start bw rows below their column).  On the
      // FP32 path Q2 runs on a row-major copy, made in place from the
      // row-major FP64 Z_T (same element size): a window of rows is then
      // contiguous in HBM (1 KB per row and workgroup instead of 256-byte
      // pieces of 128 columns); one transposition gives Q1 its column-major Z.
      const bool rows = fp32 && n * ldzt * (int64_t)sizeof(cplx<T>) <= 0x7FFFFFFF;
      L.tsVTb.ensure((size_t)B * G2 * kmax2 * 2 * bwd * bwd * sizeof(cplx<T>) * 2);
      Q2Work<T> q2{};
      q2.n = n; q2.batch = B; q2.bw = bwd; q2.kmax = kmax2; q2.m = mmax;
      q2.V2 = ts.V2; q2.tau2 = ts.tau2; q2.strideV2 = ts.strideV2; q2.strideTau2 = ts.strideTau2;
      q2.VTb = L.tsVTb.as<cplx<T>>(); q2.strideVTb = G2 * kmax2 * 2 * bwd * bwd;
      q2.Vb = q2.VTb + B * q2.strideVTb;
      if (rows) {
        launch_zt_to_c64_rows(te.zt, ldzt, te.strideZt, n, mmax, B, s);
        q2.Z = reinterpret_cast<cplx<T>*>(te.zt); q2.ldz = ldzt; q2.strideZ = te.strideZt;
        q2.rowmajor = 1;
      } else {
        launch_zt_to_z<T>(te.zt, ldzt, te.strideZt, bw.Z, bw.ldz, bw.strideZ, n, mmax, B, s);
        q2.Z = bw.Z; q2.ldz = bw.ldz; q2.strideZ = bw.strideZ;
      }
      prepare_q2<T>(q2, s);
      L.tm.begin(PH_Q2, s);
      apply_q2<T>(q2, s);
      L.tm.end(PH_Q2, s);
      if (rows)
        launch_zr_to_z(reinterpret_cast<const cplx<float>*>(te.zt), ldzt, te.strideZt,
                       reinterpret_cast<cplx<float>*>(bw.Z), bw.ldz, bw.strideZ, n, mmax, B, s);
      bw.shift = bwd;
      bw.z_ready = true;
    }
    back_transform<T>(bw, s);
    L.tm.end(PH_BACK, s);
    R.Zt = L.Z.p;

    if (fp32) {
      // ---- FP64 refinement: lambda_k = Re(z_k^H D z_k) / |z_k|^2 with D in FP64
      L.Z64.ensure((size_t)B * n * mcap * sizeof(cplx<double>));
      L.num.ensure((size_t)B * mcap * sizeof(double));
      L.den.ensure((size_t)B * mcap * sizeof(double));
      launch_fill_zero(L.num.p, B * mcap * sizeof(double), s);
      const bool forms = h.use_sparse();
      // the FP64 residual correction of the vectors (range A) forms the
      // velocity forms of the corrected vectors itself
      const bool vref = all_modes && h.vec_refine > 0 && mmax > 1;
      if (forms) {
        // Rayleigh numerators from the nonzero blocks of D, with the FP32
        // vectors in a mode-contiguous copy (exact in FP64); with velocities
        // requested (and no vector correction to follow), the three velocity
        // forms come from the same pass over the vectors (one launch, timed
        // as the velocity phase; the velocity step then only scales them)
        const bool v = h.opts.vel != 0 && !vref;
        if (v) {
          L.vacc.ensure((size_t)B * mcap * 6 * sizeof(double));
          launch_fill_zero(L.vacc.p, B * mcap * 3 * sizeof(double), s);
          L.tm.begin(PH_VEL, s);
        }
        L.Zt.ensure((size_t)B * n * mcap * sizeof(cplx<float>));
        launch_transpose<float>(L.Z.as<cplx<float>>(), n, mmax, n, n * mcap, L.Zt.as<cplx<float>>(),
                                mcap, n * mcap, B, s);
        L.sws.ensure((size_t)sparse_dots_ws_bytes(h.N, mmax, B, v ? 4 : 1));
        launch_sparse_dots<float>(h.sp_rowptr.as<int64_t>(), h.sp_col.as<int32_t>(), h.sp_nblk,
                                  h.sp_phit.as<double>(), L.bv.as<cplx<double>>(), 0, v ? 4 : 1,
                                  L.Zt.as<cplx<float>>(), mcap, n * mcap, h.N, mmax, B,
                                  L.num.as<double>(), mcap, v ? L.vacc.as<double>() : nullptr, mcap,
                                  3 * mcap, L.sws.as<double>(), s);
        if (v) L.tm.end(PH_VEL, s);
      }
      L.tm.begin(PH_REFINE, s);
      launch_convert<float, double>(L.Z.as<cplx<float>>(), L.Z64.as<cplx<double>>(), B * n * mcap, s);
      if (!forms) {
        GemmParams<double> g{};
        g.opA = 'N'; g.opB = 'N';
        g.m = n; g.n = mmax; g.k = n;
        g.alpha = cplx<double>{1, 0};
        g.A = D64; g.lda = h.ldA; g.strideA = h.ldA * n;
        g.B = L.Z64.as<cplx<double>>(); g.ldb = n; g.strideB = n * mcap;
        g.batch = B;
        g.dot_out = L.num.as<double>(); g.dot_stride = mcap;
        g.herm_dot = 1;            // D is Hermitian: lower triangle, half the work
        g.dotX = L.Z64.as<cplx<double>>(); g.lddx = n; g.strideX = n * mcap;
        L.gws.ensure((size_t)gemm_ws_bytes(g));
        g.ws = L.gws.p;
        launch_gemm<double>(g, s);
      }
      launch_colnorm2<double>(L.Z64.as<cplx<double>>(), n, mmax, n, n * mcap, B, L.den.as<double>(),
                              mcap, s);
      launch_rq_finish(L.num.as<double>(), L.den.as<double>(), te.w, B, mmax, mcap, s);
      // Rayleigh-Ritz in FP64 over each cluster of close quotients, whose
      // FP32 vectors mix inside the cluster: on the FP32 vectors, or, when
      // the vectors are corrected, on the corrected ones (an accurate span)
      const bool crr = all_modes && h.cluster_rr && mmax > 1;
      if (crr) {
        L.cmeta.ensure((size_t)B * kClusterMeta * sizeof(int32_t));
        launch_cluster_find(te.w, mmax, mcap, cluster_slots(n), L.cmeta.as<int32_t>(), B, s);
      }
      // the reduced matrices A are dead here: scratch for the steps below
      const size_t a_bytes = (size_t)B * h.ldA * n * sizeof(cplx<T>);
      if (crr && !vref) {
        if (forms)
          cluster_rayleigh_ritz<float>(h, L, B, D64, L.Zt.as<cplx<float>>(), mcap, 1,
                                       h.opts.vel != 0, te.w, mcap, A, a_bytes, s);
        else
          cluster_rayleigh_ritz<float>(h, L, B, D64, L.Z.as<cplx<float>>(), 1, n, false, te.w, mcap,
                                       A, a_bytes, s);
      }
      if (vref) {
        // the vectors are still the FP32 ones (no cluster step ran on them)
        correct_vectors(h, L, B, D64, reinterpret_cast<cplx<float>*>(A),
                        reinterpret_cast<cplx<float>*>(te.zt), te.w, mmax, mcap, true, true, s);
        // the Rayleigh quotients of the corrected vectors
        launch_rq_finish(L.num.as<double>(), L.den.as<double>(), te.w, B, mmax, mcap, s);
        if (crr)
          cluster_rayleigh_ritz<double>(h, L, B, D64, L.Z64.as<cplx<double>>(), 1, n,
                                        forms && h.opts.vel != 0, te.w, mcap, A, a_bytes, s);
      }
      if (all_modes)
        refine_low_modes_dev(h, L, B, D64, L.Z64.as<cplx<double>>(), te.w, L.den.as<double>(), mcap, s);
      else
        refine_low_modes(h, L, B, D64, L.Z64.as<cplx<double>>(), te.w, L.den.as<double>(), R.M, ilo,
                         mcap, s);
      L.tm.end(PH_REFINE, s);
      R.Z64 = L.Z64.as<cplx<double>>();
      if (o.range == 'V') {
        // RANGE = 'V' selects lambda in (vl, vu] by Sturm counts of the FP32
        // tridiagonal; a refined eigenvalue that left the interval (one within
        // the FP32 error of a bound) is dropped.  (A mode that the FP32 counts
        // put just outside is not computed, so it cannot be added.)
        std::vector<double> wl((size_t)B * mcap);
        PPH_HIP(hipMemcpyAsync(wl.data(), te.w, wl.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        PPH_HIP(hipStreamSynchronize(s));
        for (int64_t b = 0; b < B; ++b) {
          const double* wb = wl.data() + b * mcap;
          int64_t f = 0, l = R.M[b];
          while (f < l && !(wb[f] > vl_l)) ++f;
          while (l > f && wb[l - 1] > vu_l) --l;
          R.first[b] = f;
          R.M[b] = l - f;
        }
      }
    } else {
      R.Z64 = reinterpret_cast<cplx<double>*>(L.Z.p);
    }
  }
}

// ---------------------------------------------------------------- run one batch
struct BatchOut {
  EigResult R;
  double* om = nullptr;        // device B x mcap, cm^-1
  double* vel = nullptr;       // device B x mcap x 3
};

// Everything of one batch after its q-points are in L.qdev: only kernels
// on the lane's stream (and, outside range A, the host reads of the mode
// counts), so a range-A batch can be captured as a graph.
static void compute_batch(Handle& h, Lane& L, int64_t B, BatchOut& out) {
  const int64_t n = h.n, N = h.N;
  hipStream_t s = L.stream;
  const pph_opts& o = h.opts;
  const bool fp32 = (o.precision == 'S');
  const bool vel = o.vel != 0;
  const bool want_vec = vel || o.jobz == 'V';
  const bool sparse = h.use_sparse();
  // D in FP64: the FP64 path's working matrix; on the FP32 path the
  // refinement's dense products read it, which the block-sparse forms replace
  const bool d64 = !fp32 || !sparse;
  if (d64) L.A64.ensure((size_t)B * h.ldA * n * sizeof(cplx<double>));
  if (fp32) L.A32.ensure((size_t)B * h.ldA * n * sizeof(cplx<float>));
  if (vel && !sparse)
    L.dD.ensure((size_t)B * 3 * n * n * (fp32 ? sizeof(cplx<float>) : sizeof(cplx<double>)));

  AsmParams ap{};
  ap.phi = h.phi.as<double>(); ap.crys = h.crys.as<double>(); ap.lat = h.lat.as<double>();
  ap.rmass = h.rmass.as<double>(); ap.mask = h.mask.as<uint16_t>(); ap.q = L.qdev.as<double>();
  ap.cd_dc = h.cd_dc.as<double>(); ap.cd_norm = h.cd_norm;
  ap.N = N; ap.nb = B; ap.mode = (o.vel_method == 'C') ? 1 : 0;
  ap.D64 = d64 ? L.A64.as<cplx<double>>() : nullptr;
  ap.D32 = fp32 ? L.A32.as<cplx<float>>() : nullptr;
  ap.ldD = h.ldA;
  ap.nzflag = h.sp_flags.as<uint8_t>();
  if (vel && !sparse) {
    if (fp32) ap.dD32 = L.dD.as<cplx<float>>();
    else ap.dD64 = L.dD.as<cplx<double>>();
  }
  L.tm.begin(PH_TOTAL, s);
  L.tm.begin(PH_ASM, s);
  launch_assemble(ap, s);
  if (sparse) {
    // phase sums of D (for the FP32 path's Rayleigh quotients) and dD on the
    // nonzero blocks
    L.bv.ensure((size_t)B * h.sp_nblk * 4 * sizeof(cplx<double>));
    launch_sparse_blocks(ap, h.sp_row.as<int32_t>(), h.sp_col.as<int32_t>(), h.sp_nblk, vel ? 4 : 1,
                         L.bv.as<cplx<double>>(), s);
  }
  L.tm.end(PH_ASM, s);

  const double K32 = (double)(float)(15.633302f * 33.35641f);
  const double vl_l = o.vl / K32, vu_l = o.vu / K32;
  EigResult& R = out.R;
  if (fp32)
    eig_batch<float>(h, L, B, L.A32.as<cplx<float>>(), d64 ? L.A64.as<cplx<double>>() : nullptr,
                     want_vec, vl_l, vu_l, R);
  else eig_batch<double>(h, L, B, L.A64.as<cplx<double>>(), nullptr, want_vec, vl_l, vu_l, R);

  const int64_t mcap = R.mcap;
  int64_t mmax = 0;    // modes 0 .. first + M - 1 of each matrix
  for (int64_t b = 0; b < B; ++b) mmax = std::max(mmax, R.first[b] + R.M[b]);
  L.om.ensure((size_t)B * mcap * sizeof(double));
  launch_eval_to_cm(R.lam, L.om.as<double>(), B, mmax, mcap, s);
  out.om = L.om.as<double>();

  // the FP32 path with block-sparse forms formed the velocity forms in the
  // same launch as its Rayleigh numerators (timed there as the velocity phase)
  const bool vel_done = sparse && fp32;
  if (vel && mmax > 0 && !vel_done) {
    L.vacc.ensure((size_t)B * mcap * 6 * sizeof(double));
    launch_fill_zero(L.vacc.p, B * mcap * 3 * sizeof(double), s);
    L.tm.begin(PH_VEL, s);
    if (sparse) {
      L.Zt.ensure((size_t)B * n * mcap * sizeof(cplx<double>));
      launch_transpose<double>(L.Z.as<cplx<double>>(), n, mmax, n, n * mcap,
                               L.Zt.as<cplx<double>>(), mcap, n * mcap, B, s);
      L.sws.ensure((size_t)sparse_dots_ws_bytes(N, mmax, B, 3));
      launch_sparse_dots<double>(h.sp_rowptr.as<int64_t>(), h.sp_col.as<int32_t>(), h.sp_nblk,
                                 h.sp_phit.as<double>(), L.bv.as<cplx<double>>(), 1, 3,
                                 L.Zt.as<cplx<double>>(), mcap, n * mcap, N, mmax, B, nullptr, 0,
                                 L.vacc.as<double>(), mcap, 3 * mcap, L.sws.as<double>(), s);
    } else if (fp32) {
      GemmParams<float> g{};
      g.opA = 'N'; g.opB = 'N';
      g.m = n; g.n = mmax; g.k = n;
      g.alpha = cplx<float>{1, 0};
      g.A = L.dD.as<cplx<float>>(); g.lda = n; g.strideA = 3 * n * n; g.sAin = n * n;
      g.B = L.Z.as<cplx<float>>(); g.ldb = n; g.strideB = n * mcap; g.sBin = 0;
      g.batch = B; g.binner = 3;
      g.dot_out = L.vacc.as<double>(); g.dot_stride = 3 * mcap; g.sDotin = mcap;
      g.herm_dot = 1;            // dD/dq is Hermitian: lower triangle, half the work
      g.dotX = L.Z.as<cplx<float>>(); g.lddx = n; g.strideX = n * mcap; g.sXin = 0;
      L.gws.ensure((size_t)gemm_ws_bytes(g));
      g.ws = L.gws.p;
      launch_gemm<float>(g, s);
    } else {
      GemmParams<double> g{};
      g.opA = 'N'; g.opB = 'N';
      g.m = n; g.n = mmax; g.k = n;
      g.alpha = cplx<double>{1, 0};
      g.A = L.dD.as<cplx<double>>(); g.lda = n; g.strideA = 3 * n * n; g.sAin = n * n;
      g.B = L.Z.as<cplx<double>>(); g.ldb = n; g.strideB = n * mcap; g.sBin = 0;
      g.batch = B; g.binner = 3;
      g.dot_out = L.vacc.as<double>(); g.dot_stride = 3 * mcap; g.sDotin = mcap;
      g.herm_dot = 1;            // dD/dq is Hermitian: lower triangle, half the work
      g.dotX = L.Z.as<cplx<double>>(); g.lddx = n; g.strideX = n * mcap; g.sXin = 0;
      L.gws.ensure((size_t)gemm_ws_bytes(g));
      g.ws = L.gws.p;
      launch_gemm<double>(g, s);
    }
    L.tm.end(PH_VEL, s);
  }
  if (vel && mmax > 0) {
    launch_vel_scale(L.vacc.as<double>(), L.om.as<double>(), B, mmax, mcap, mcap, s);
    out.vel = L.vacc.as<double>() + B * 3 * mcap;
  }
  L.tm.end(PH_TOTAL, s);
}

// One batch: the q-points to the device, then compute_batch, eagerly or as a
// HIP graph.  A range-A batch of a shape the lane has just run eagerly (so
// every buffer is allocated at its final size) is captured once and then
// replayed for every later batch of that shape while no device buffer has
// been allocated or freed since the capture; the captured batch's host-side
// results (counts, device pointers) are those of every replay.  Profiling,
// ranges I and V, and a failed capture run eagerly.
static void run_batch(Handle& h, Lane& L, int64_t B, const double* q_host_rows /* B x 3 row-major */,
                      BatchOut& out) {
  hipStream_t s = L.stream;
  L.qdev.ensure((size_t)B * 3 * sizeof(double));
  PPH_HIP(hipMemcpyAsync(L.qdev.p, q_host_rows, B * 3 * sizeof(double), hipMemcpyHostToDevice, s));
  const bool eligible = h.graphs && !L.graph_failed && h.opts.range == 'A' && !L.tm.on;
  const std::array<int64_t, 8> key{B, h.nb_panel, h.nb_back, (h.use_sparse() ? 1 : 0) | (h.cluster_rr ? 2 : 0) | (h.vec_refine << 2) |
                                       (h.vec_bf16 ? 16 : 0),
                                   h.reduction_code(), h.band_width(), h.n,
                                   h.gemm_tile ? h.gemm_tile : gemm_current_tile()};
  if (eligible && L.graph.exec && L.graph.key == key && L.graph.gen == g_alloc_gen.load() &&
      L.graph_out) {
    PPH_HIP(hipGraphLaunch(L.graph.exec, s));
    out = *L.graph_out;
    ++L.graph_replays;
    return;
  }
  const uint64_t gen0 = g_alloc_gen.load();
  if (eligible && L.eager_key == key && L.eager_gen == gen0) {
    if (L.graph.exec) {
      PPH_HIP(hipGraphExecDestroy(L.graph.exec));
      L.graph.exec = nullptr;
    }
    PPH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipGraph_t g = nullptr;
    bool ok = true;
    try {
      compute_batch(h, L, B, out);
    } catch (...) {
      ok = false;
    }
    const hipError_t e = hipStreamEndCapture(s, &g);
    ok = ok && e == hipSuccess && g != nullptr;
    // another lane allocating meanwhile does not make the capture wrong, but
    // the count it would be replayed against is then unknown: run eagerly
    const bool same_gen = g_alloc_gen.load() == gen0;
    if (ok && same_gen) ok = hipGraphInstantiate(&L.graph.exec, g, nullptr, nullptr, 0) == hipSuccess;
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    if (ok && same_gen) {
      L.graph.key = key;
      L.graph.gen = gen0;
      L.graph_out = std::make_shared<BatchOut>(out);
      ++L.graph_captures;
      PPH_HIP(hipGraphLaunch(L.graph.exec, s));
      return;
    }
    L.graph.exec = nullptr;
    if (!ok) L.graph_failed = true;      // not capturable: this lane stays eager
  }
  compute_batch(h, L, B, out);
  L.eager_key = key;
  L.eager_gen = g_alloc_gen.load();
}

// q-points per device batch when the caller does not fix it: as many as fit
// in half of the device's free memory shared by `lanes` batches in flight
// (larger batches keep more of the
// per-column launches of the reduction busy), at most 64.  Per matrix: D in
// FP64 (16 n^2) and FP32 (8 n^2), dD (48 or 96 n^2), Householder vectors,
// eigenvectors (FP32 + FP64 copies, the FP64 tridiagonal vectors and their
// LU factors), about 6 n^2 x 16 bytes more; with block-sparse forms no dD but
// a mode-contiguous copy of the vectors.
static int64_t auto_batch(const Handle& h, int64_t nq, int64_t lanes) {
  if (h.opts.batch > 0) return std::min(nq, h.opts.batch);
  const int64_t n = h.n;
  const bool s = h.opts.precision == 'S';
  const bool sp = h.use_sparse();
  const double per = (double)n * n * ((s && sp ? 0.0 : 16.0) + (s ? 8.0 : 0.0) +
                                      (h.opts.vel && !sp ? (s ? 48.0 : 96.0) : 0.0) +
                                      (sp ? (s ? 8.0 : 16.0) : 0.0) + 96.0);
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)24e9;
  int64_t b = (int64_t)std::floor(0.5 * (double)free_b / std::max<int64_t>(1, lanes) /
                                  std::max(per, 1.0));
  b = std::max<int64_t>(1, std::min<int64_t>(b, 64));
  return std::min(nq, b);
}

}  // namespace pph

using namespace pph;

static int fail(Handle* h, const std::string& msg, int code) {
  if (h) h->err = msg;
  g_last_error = msg;
  return code;
}

#define PPH_API_TRY try {
#define PPH_API_CATCH(h)                                                     \
  }                                                                          \
  catch (const std::bad_alloc&) {                                            \
    return fail(h, "out of device or host memory", PPH_ERR_MEMORY);          \
  }                                                                          \
  catch (const pph::hip_error& e) {                                          \
    return fail(h, e.what(), PPH_ERR_HIP);                                   \
  }                                                                          \
  catch (const std::exception& e) {                                          \
    return fail(h, e.what(), PPH_ERR_INTERNAL);                              \
  }

static int check_opts(const pph_opts* o, std::string& why) {
  if (!o) { why = "opts is NULL"; return -1; }
  if (o->jobz != 'N' && o->jobz != 'V') { why = "jobz must be 'N' or 'V'"; return -1; }
  if (o->range != 'A' && o->range != 'V' && o->range != 'I') { why = "range must be 'A', 'V' or 'I'"; return -1; }
  if (o->vel_method != 'A' && o->vel_method != 'C') { why = "vel_method must be 'A' or 'C'"; return -1; }
  if (o->precision != 'D' && o->precision != 'S') { why = "precision must be 'D' or 'S'"; return -1; }
  if (o->range == 'I' && o->il > o->iu) { why = "range I needs il <= iu"; return -1; }
  if (o->range == 'V' && o->vl > o->vu) { why = "range V needs vl <= vu"; return -1; }
  if (o->range == 'I' && o->il < 1) { why = "range I needs il >= 1"; return -1; }
  if (o->streams < 0 || o->streams > 8) { why = "streams must be 0..8"; return -1; }
  return 0;
}

extern "C" {

const char* pph_version(void) { return "parphom_hip 0.1 (gfx950)"; }

const char* pph_last_error(const pph_handle* h) {
  if (h) return reinterpret_cast<const Handle*>(h)->err.c_str();
  return g_last_error.c_str();
}

// Nonzero 3x3 blocks of Phi_sym, from the device copy.  The lower lists
// (a >= b, the diagonal always) are kept when at most a quarter of all N^2
// blocks are nonzero, or for N <= 2048 (so that the block-sparse path can be
// selected for testing).
static void build_block_lists(Handle& h) {
  const int64_t N = h.N;
  hipStream_t s = h.stream;
  std::vector<uint8_t> flags((size_t)N * N);
  // the device flags stay with the handle: the assembly skips the phase sums
  // of zero blocks
  h.sp_flags.ensure((size_t)N * N);
  launch_block_nz(h.phi.as<double>(), h.sp_flags.as<uint8_t>(), N, s);
  PPH_HIP(hipMemcpyAsync(flags.data(), h.sp_flags.p, (size_t)N * N, hipMemcpyDeviceToHost, s));
  PPH_HIP(hipStreamSynchronize(s));
  int64_t full = 0;                 // nonzero blocks of the whole matrix
  for (size_t i = 0; i < flags.size(); ++i) full += flags[i];
  const int64_t cap = N * N / 4;
  if (full > cap && N > 2048) return;
  std::vector<int64_t> rowptr(N + 1, 0);
  std::vector<int32_t> row, col;
  for (int64_t a = 0; a < N; ++a) {
    const uint8_t* fa = flags.data() + (size_t)a * N;      // flags(b, a) = flags(a, b)
    for (int64_t b = 0; b <= a; ++b)
      if (b == a || fa[b]) {
        row.push_back((int32_t)a);
        col.push_back((int32_t)b);
      }
    rowptr[a + 1] = (int64_t)row.size();
  }
  // transposed lists: for atom b, the blocks (a, b) with a > b
  std::vector<int64_t> tptr(N + 1, 0), tblk;
  std::vector<int32_t> trow;
  for (size_t k = 0; k < row.size(); ++k)
    if (col[k] < row[k]) ++tptr[col[k] + 1];
  for (int64_t a = 0; a < N; ++a) tptr[a + 1] += tptr[a];
  trow.resize(tptr[N]);
  tblk.resize(tptr[N]);
  {
    std::vector<int64_t> fill(tptr.begin(), tptr.end() - 1);
    for (size_t k = 0; k < row.size(); ++k)
      if (col[k] < row[k]) {
        const int64_t t = fill[col[k]]++;
        trow[t] = row[k];
        tblk[t] = (int64_t)k;
      }
  }
  h.sp_ok = full <= cap;
  h.sp_nblk = (int64_t)row.size();
  const int64_t nt = tptr[N];
  h.sp_tptr.ensure((N + 1) * sizeof(int64_t));
  h.sp_trow.ensure(std::max<int64_t>(nt, 1) * sizeof(int32_t));
  h.sp_tblk.ensure(std::max<int64_t>(nt, 1) * sizeof(int64_t));
  PPH_HIP(hipMemcpyAsync(h.sp_tptr.p, tptr.data(), (N + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
  if (nt > 0) {
    PPH_HIP(hipMemcpyAsync(h.sp_trow.p, trow.data(), nt * sizeof(int32_t), hipMemcpyHostToDevice, s));
    PPH_HIP(hipMemcpyAsync(h.sp_tblk.p, tblk.data(), nt * sizeof(int64_t), hipMemcpyHostToDevice, s));
  }
  h.sp_rowptr.ensure((N + 1) * sizeof(int64_t));
  h.sp_row.ensure(h.sp_nblk * sizeof(int32_t));
  h.sp_col.ensure(h.sp_nblk * sizeof(int32_t));
  PPH_HIP(hipMemcpyAsync(h.sp_rowptr.p, rowptr.data(), (N + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
  PPH_HIP(hipMemcpyAsync(h.sp_row.p, row.data(), h.sp_nblk * sizeof(int32_t), hipMemcpyHostToDevice, s));
  PPH_HIP(hipMemcpyAsync(h.sp_col.p, col.data(), h.sp_nblk * sizeof(int32_t), hipMemcpyHostToDevice, s));
  h.sp_phit.ensure(h.sp_nblk * 9 * sizeof(double));
  launch_block_phi(h.phi.as<double>(), h.mask.as<uint16_t>(), h.rmass.as<double>(),
                   h.sp_row.as<int32_t>(), h.sp_col.as<int32_t>(), h.sp_nblk, N,
                   h.sp_phit.as<double>(), s);
  PPH_HIP(hipStreamSynchronize(s));
}

int pph_create(const pph_system* sys, const double* fc, const pph_opts* opts, pph_handle** out) {
  if (!out) return fail(nullptr, "pph_create: out is NULL", PPH_ERR_ARG);
  *out = nullptr;
  if (!sys || !fc) return fail(nullptr, "pph_create: sys or fc is NULL", PPH_ERR_ARG);
  if (sys->natom < 1 || sys->ntypes < 1 || !sys->lat || !sys->pos || !sys->mass || !sys->type)
    return fail(nullptr, "pph_create: incomplete system description", PPH_ERR_ARG);
  std::string why;
  if (check_opts(opts, why)) return fail(nullptr, "pph_create: " + why, PPH_ERR_ARG);
  for (int64_t i = 0; i < sys->natom; ++i)
    if (sys->type[i] < 1 || sys->type[i] > sys->ntypes)
      return fail(nullptr, "pph_create: atom type out of range", PPH_ERR_ARG);
  std::unique_ptr<Handle> h(new Handle());
  PPH_API_TRY
  h->opts = *opts;
  if (h->opts.vel) h->opts.jobz = 'V';
  h->device = opts->device;
  PPH_HIP(hipSetDevice(h->device));
  PPH_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  const int64_t N = sys->natom, n = 3 * N;
  h->N = N; h->set_n(n);
  Geometry& g = h->g;
  g.N = N;
  std::memcpy(g.lat, sys->lat, sizeof(g.lat));
  g.pos.assign(sys->pos, sys->pos + 3 * N);
  g.crys.resize(3 * N);
  for (int64_t i = 0; i < N; ++i) {
    const double x[3] = {g.pos[i], g.pos[i + N], g.pos[i + 2 * N]};
    double u[3];
    linsolve(x, g.lat, u);
    g.crys[i] = u[0]; g.crys[i + N] = u[1]; g.crys[i + 2 * N] = u[2];
  }
  // rec = 2 pi inv(lat)^T (matinvtransp3)
  {
    double Binv[9];
    inv3(g.lat, Binv, 6.283185307179586);
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) g.rec[i + 3 * k] = Binv[k + 3 * i];
  }
  g.rmass.resize(N);
  for (int64_t i = 0; i < N; ++i) g.rmass[i] = sys->mass[sys->type[i] - 1];
  // central-difference steps (dynamical_matrix.F90:136-142)
  const double dinc = (double)1e-6f;
  const double sdel = std::max(dinc * std::sqrt(std::fabs(L(g.rec, 0, 0) * L(g.rec, 1, 1) -
                                                          L(g.rec, 0, 1) * L(g.rec, 1, 0))), dinc);
  double dc[9];
  for (int a = 0; a < 3; ++a) {
    double delta[3] = {0, 0, 0};
    delta[a] = sdel;
    double u[3];
    linsolve(delta, g.rec, u);
    for (int c = 0; c < 3; ++c) dc[3 * a + c] = u[c];
  }
  h->cd_norm = sdel;

  hipStream_t s = h->stream;
  h->phi.ensure((size_t)n * n * sizeof(double));
  {
    DevBuf fcd;
    fcd.ensure((size_t)N * N * 9 * sizeof(double));
    PPH_HIP(hipMemcpyAsync(fcd.p, fc, (size_t)N * N * 9 * sizeof(double), hipMemcpyHostToDevice, s));
    launch_symmetrise_fc(fcd.as<double>(), h->phi.as<double>(), N, s);
    PPH_HIP(hipStreamSynchronize(s));
  }
  h->crys.ensure(3 * N * sizeof(double));
  h->pos.ensure(3 * N * sizeof(double));
  h->lat.ensure(9 * sizeof(double));
  h->rmass.ensure(N * sizeof(double));
  h->cd_dc.ensure(9 * sizeof(double));
  h->mask.ensure((size_t)N * N * sizeof(uint16_t));
  PPH_HIP(hipMemcpyAsync(h->crys.p, g.crys.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice, s));
  PPH_HIP(hipMemcpyAsync(h->pos.p, g.pos.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice, s));
  PPH_HIP(hipMemcpyAsync(h->lat.p, g.lat, 9 * sizeof(double), hipMemcpyHostToDevice, s));
  PPH_HIP(hipMemcpyAsync(h->rmass.p, g.rmass.data(), N * sizeof(double), hipMemcpyHostToDevice, s));
  PPH_HIP(hipMemcpyAsync(h->cd_dc.p, dc, 9 * sizeof(double), hipMemcpyHostToDevice, s));
  launch_image_table(h->crys.as<double>(), h->pos.as<double>(), h->lat.as<double>(),
                     h->mask.as<uint16_t>(), N, s);
  build_block_lists(*h);
  PPH_HIP(hipStreamSynchronize(s));
  *out = reinterpret_cast<pph_handle*>(h.release());
  return PPH_OK;
  PPH_API_CATCH(nullptr)
}

static int run_common(pph_handle* hh, int64_t nq, const double* q, int64_t ldq, pph_out* out,
                      bool device_out) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h) return fail(nullptr, "pph_run: handle is NULL", PPH_ERR_ARG);
  if (nq < 0 || (nq > 0 && (!q || ldq < nq))) return fail(h, "pph_run: bad q array", PPH_ERR_ARG);
  if (!out || !out->eval_cm || !out->m || !out->info)
    return fail(h, "pph_run: eval_cm, m and info are required", PPH_ERR_ARG);
  const int64_t n = h->n;
  const int64_t mcap = range_mcap(h->opts, n);
  if (out->ld_eval < mcap) return fail(h, "pph_run: ld_eval smaller than the eigenvalue count", PPH_ERR_ARG);
  if ((out->evec || out->vel) && out->evec_cols <= 0)
    return fail(h, "pph_run: evec_cols must be > 0 with evec or vel", PPH_ERR_ARG);
  if (out->vel && !h->opts.vel) return fail(h, "pph_run: vel requested but the handle has vel = 0", PPH_ERR_ARG);
  if (out->evec && h->opts.jobz != 'V')
    return fail(h, "pph_run: evec requested but jobz = 'N'", PPH_ERR_ARG);
  PPH_API_TRY
  PPH_HIP(hipSetDevice(h->device));
  const hipMemcpyKind kind = device_out ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  // lanes: opts.streams, or two when the q-points make at least two batches
  const int64_t want = h->opts.streams > 0 ? h->opts.streams : 2;
  int64_t bmax = auto_batch(*h, std::max<int64_t>(nq, 1), want);
  const int64_t nbatch = ceil_div(std::max<int64_t>(nq, 1), bmax);
  const int64_t nl = std::max<int64_t>(1, std::min<int64_t>(want, nbatch));
  if (nl == 1 && h->opts.batch <= 0) bmax = auto_batch(*h, std::max<int64_t>(nq, 1), 1);
  while ((int64_t)h->lanes() < nl) {
    std::unique_ptr<Lane> ln(new Lane());
    // a lane added while profiling is on times its batches like lane 0
    ln->tm.on = h->tm.on;
    if (ln->tm.on) ln->tm.init();
    PPH_HIP(hipStreamCreateWithFlags(&ln->stream, hipStreamNonBlocking));
    h->more.push_back(std::move(ln));
  }
  // lane l runs batches l, l + nl, ... and copies their outputs on its own stream
  auto work = [&](int64_t l) {
    PPH_HIP(hipSetDevice(h->device));
    Lane& L = h->lane((size_t)l);
    const hipStream_t s = L.stream;
    // allocated before the first batch, so a later batch can be captured
    if (device_out) L.meta.ensure((size_t)6 * bmax * sizeof(int64_t));
    std::vector<double> qrows;
    for (int64_t q0 = l * bmax; q0 < nq; q0 += nl * bmax) {
      const int64_t B = std::min(bmax, nq - q0);
      qrows.resize(3 * B);
      for (int64_t b = 0; b < B; ++b)
        for (int c = 0; c < 3; ++c) qrows[3 * b + c] = q[(q0 + b) + ldq * c];
      BatchOut bo;
      run_batch(*h, L, B, qrows.data(), bo);
      L.tm.collect();
      const EigResult& R = bo.R;
      std::vector<int32_t> infos(B);
      std::vector<int64_t> nzs(B);
      // per matrix, in doubles: offset and count of the eigenvalues, the
      // eigenvectors and the velocities to return
      std::vector<int64_t> meta(6 * B, 0);
      int64_t mmax = 0, nzmax = 0;
      for (int64_t b = 0; b < B; ++b) {
        const int64_t m = R.M[b];
        int32_t info = R.info[b];
        int64_t nz = 0;
        if (h->opts.jobz == 'V') {
          nz = m;
          if (out->evec || out->vel) {
            if (nz > out->evec_cols) { info |= PPH_INFO_MEMORY; nz = out->evec_cols; }
          }
        }
        infos[b] = info;
        nzs[b] = nz;
        const int64_t f0 = b * R.mcap + R.first[b];      // first returned mode
        meta[b] = f0;                 meta[B + b] = m;
        meta[2 * B + b] = f0 * n * 2; meta[3 * B + b] = nz * n * 2;
        meta[4 * B + b] = f0 * 3;     meta[5 * B + b] = nz * 3;
        mmax = std::max(mmax, m);
        nzmax = std::max(nzmax, nz);
      }
      std::vector<int32_t> dflags;
      if (device_out) {
        // one gather launch per output array for the whole batch
        L.meta.ensure(meta.size() * sizeof(int64_t));
        int64_t* md = L.meta.as<int64_t>();
        PPH_HIP(hipMemcpyAsync(md, meta.data(), meta.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
        launch_gather_rows(bo.om, md, md + B, out->eval_cm + q0 * out->ld_eval, out->ld_eval, mmax, B, s);
        if (out->evec && nzmax > 0)
          launch_gather_rows(reinterpret_cast<const double*>(R.Z64), md + 2 * B, md + 3 * B,
                             reinterpret_cast<double*>(out->evec) + q0 * n * out->evec_cols * 2,
                             n * out->evec_cols * 2, nzmax * n * 2, B, s);
        if (out->vel && nzmax > 0)
          launch_gather_rows(bo.vel, md + 4 * B, md + 5 * B, out->vel + q0 * out->evec_cols * 3,
                             out->evec_cols * 3, nzmax * 3, B, s);
        PPH_HIP(hipMemcpyAsync(out->m + q0, R.M.data(), B * sizeof(int64_t), hipMemcpyHostToDevice, s));
        PPH_HIP(hipMemcpyAsync(out->info + q0, infos.data(), B * sizeof(int32_t), hipMemcpyHostToDevice, s));
        if (R.info_dev) launch_or_flags(out->info + q0, R.info_dev, B, s);
        if (out->nz) PPH_HIP(hipMemcpyAsync(out->nz + q0, nzs.data(), B * sizeof(int64_t), hipMemcpyHostToDevice, s));
      } else {
        for (int64_t b = 0; b < B; ++b) {
          const int64_t iq = q0 + b, f0 = meta[b], m = meta[B + b], nz = nzs[b];
          if (m > 0)
            PPH_HIP(hipMemcpyAsync(out->eval_cm + iq * out->ld_eval, bo.om + f0, m * sizeof(double),
                                   kind, s));
          if (out->evec && nz > 0)
            PPH_HIP(hipMemcpyAsync((char*)out->evec + (size_t)iq * n * out->evec_cols * 16,
                                   R.Z64 + (size_t)f0 * n, (size_t)nz * n * 16, kind, s));
          if (out->vel && nz > 0)
            PPH_HIP(hipMemcpyAsync(out->vel + (size_t)iq * out->evec_cols * 3, bo.vel + f0 * 3,
                                   (size_t)nz * 3 * sizeof(double), kind, s));
        }
        if (R.info_dev) {
          dflags.assign(B, 0);
          PPH_HIP(hipMemcpyAsync(dflags.data(), R.info_dev, B * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        }
      }
      PPH_HIP(hipStreamSynchronize(s));      // the host vectors above outlive the copies
      if (!device_out) {
        for (int64_t b = 0; b < B; ++b) {
          out->m[q0 + b] = R.M[b];
          out->info[q0 + b] = infos[b] | (dflags.empty() ? 0 : dflags[b]);
          if (out->nz) out->nz[q0 + b] = nzs[b];
        }
      }
    }
  };
  if (nl == 1) {
    work(0);
  } else {
    std::vector<std::exception_ptr> errs((size_t)nl);
    std::vector<std::thread> th;
    for (int64_t l = 1; l < nl; ++l)
      th.emplace_back([&, l] {
        try { work(l); } catch (...) { errs[(size_t)l] = std::current_exception(); }
      });
    try { work(0); } catch (...) { errs[0] = std::current_exception(); }
    for (auto& t : th) t.join();
    for (auto& e : errs)
      if (e) std::rethrow_exception(e);
  }
  return PPH_OK;
  PPH_API_CATCH(h)
}

int pph_run(pph_handle* h, int64_t nq, const double* q, int64_t ldq, pph_out* out) {
  return run_common(h, nq, q, ldq, out, false);
}

int pph_run_device(pph_handle* h, int64_t nq, const double* q, int64_t ldq, pph_out* out) {
  return run_common(h, nq, q, ldq, out, true);
}

int pph_assemble(pph_handle* hh, int64_t nq, const double* q, int64_t ldq, int32_t what, void* out,
                 int32_t on_device) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h) return fail(nullptr, "pph_assemble: handle is NULL", PPH_ERR_ARG);
  if (what < 0 || what > 3) return fail(h, "pph_assemble: what must be 0..3", PPH_ERR_ARG);
  if (nq < 0 || (nq > 0 && (!q || !out || ldq < nq))) return fail(h, "pph_assemble: bad arguments", PPH_ERR_ARG);
  PPH_API_TRY
  PPH_HIP(hipSetDevice(h->device));
  hipStream_t s = h->stream;
  const int64_t n = h->n;
  const int64_t bmax = std::max<int64_t>(1, std::min<int64_t>(nq, 8));
  std::vector<double> qrows;
  for (int64_t q0 = 0; q0 < nq; q0 += bmax) {
    const int64_t B = std::min(bmax, nq - q0);
    qrows.resize(3 * B);
    for (int64_t b = 0; b < B; ++b)
      for (int c = 0; c < 3; ++c) qrows[3 * b + c] = q[(q0 + b) + ldq * c];
    h->qdev.ensure((size_t)B * 3 * sizeof(double));
    PPH_HIP(hipMemcpyAsync(h->qdev.p, qrows.data(), B * 3 * sizeof(double), hipMemcpyHostToDevice, s));
    AsmParams ap{};
    ap.phi = h->phi.as<double>(); ap.crys = h->crys.as<double>(); ap.lat = h->lat.as<double>();
    ap.rmass = h->rmass.as<double>(); ap.mask = h->mask.as<uint16_t>(); ap.q = h->qdev.as<double>();
    ap.cd_dc = h->cd_dc.as<double>(); ap.cd_norm = h->cd_norm;
    ap.N = h->N; ap.nb = B; ap.mode = (h->opts.vel_method == 'C') ? 1 : 0;
    const int64_t ldA = h->ldA;
    h->A64.ensure((size_t)B * ldA * n * sizeof(cplx<double>));
    ap.D64 = h->A64.as<cplx<double>>();
    ap.ldD = ldA;
    ap.nzflag = h->sp_flags.as<uint8_t>();
    if (what != 0) {
      h->dD.ensure((size_t)B * 3 * n * n * sizeof(cplx<double>));
      ap.dD64 = h->dD.as<cplx<double>>();
    }
    launch_assemble(ap, s);
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    for (int64_t b = 0; b < B; ++b) {
      char* o = (char*)out + (size_t)(q0 + b) * n * n * 16;
      if (what == 0)   // D is stored with leading dimension ldA
        PPH_HIP(hipMemcpy2DAsync(o, n * 16, ap.D64 + (size_t)b * ldA * n, ldA * 16, n * 16, n, kind, s));
      else
        PPH_HIP(hipMemcpyAsync(o, ap.dD64 + (size_t)(3 * b + (what - 1)) * n * n, (size_t)n * n * 16,
                               kind, s));
    }
    PPH_HIP(hipStreamSynchronize(s));
  }
  return PPH_OK;
  PPH_API_CATCH(h)
}

int pph_expect(pph_handle* hh, int64_t nq, const double* q, int64_t ldq, int32_t what,
               const void* Z, int64_t nz, int64_t zcols, double* out, int32_t on_device) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h) return fail(nullptr, "pph_expect: handle is NULL", PPH_ERR_ARG);
  if (what < 0 || what > 3) return fail(h, "pph_expect: what must be 0..3", PPH_ERR_ARG);
  if (nq < 0 || nz < 0 || zcols < nz || (nq > 0 && (!q || !Z || !out || ldq < nq)))
    return fail(h, "pph_expect: bad arguments", PPH_ERR_ARG);
  PPH_API_TRY
  PPH_HIP(hipSetDevice(h->device));
  hipStream_t s = h->stream;
  const int64_t n = h->n;
  const int64_t B = std::max<int64_t>(1, std::min<int64_t>(nq, 4));
  DevBuf zb, ob, wsb;
  std::vector<double> qrows;
  for (int64_t q0 = 0; q0 < nq; q0 += B) {
    const int64_t nb = std::min(B, nq - q0);
    qrows.resize(3 * nb);
    for (int64_t b = 0; b < nb; ++b)
      for (int c = 0; c < 3; ++c) qrows[3 * b + c] = q[(q0 + b) + ldq * c];
    h->qdev.ensure((size_t)nb * 3 * sizeof(double));
    PPH_HIP(hipMemcpyAsync(h->qdev.p, qrows.data(), nb * 3 * sizeof(double), hipMemcpyHostToDevice, s));
    AsmParams ap{};
    ap.phi = h->phi.as<double>(); ap.crys = h->crys.as<double>(); ap.lat = h->lat.as<double>();
    ap.rmass = h->rmass.as<double>(); ap.mask = h->mask.as<uint16_t>(); ap.q = h->qdev.as<double>();
    ap.cd_dc = h->cd_dc.as<double>(); ap.cd_norm = h->cd_norm;
    ap.N = h->N; ap.nb = nb; ap.mode = (h->opts.vel_method == 'C') ? 1 : 0;
    h->A64.ensure((size_t)nb * h->ldA * n * sizeof(cplx<double>));
    ap.D64 = h->A64.as<cplx<double>>();
    ap.ldD = h->ldA;
    ap.nzflag = h->sp_flags.as<uint8_t>();
    if (what > 0) {
      h->dD.ensure((size_t)nb * 3 * n * n * sizeof(cplx<double>));
      ap.dD64 = h->dD.as<cplx<double>>();
    }
    launch_assemble(ap, s);
    const cplx<double>* Zq = reinterpret_cast<const cplx<double>*>(Z) + (size_t)q0 * n * zcols;
    if (!on_device) {
      zb.ensure((size_t)nb * n * zcols * sizeof(cplx<double>));
      PPH_HIP(hipMemcpyAsync(zb.p, Zq, (size_t)nb * n * zcols * 16, hipMemcpyHostToDevice, s));
      Zq = zb.as<cplx<double>>();
    }
    double* od = on_device ? out + q0 * nz : nullptr;
    if (!on_device) {
      ob.ensure((size_t)nb * nz * sizeof(double));
      od = ob.as<double>();
    }
    launch_fill_zero(od, nb * nz * sizeof(double), s);
    GemmParams<double> g{};
    g.opA = 'N'; g.opB = 'N'; g.m = n; g.n = nz; g.k = n;
    g.alpha = cplx<double>{1, 0};
    g.A = (what == 0) ? h->A64.as<cplx<double>>() : h->dD.as<cplx<double>>() + (size_t)(what - 1) * n * n;
    g.lda = (what == 0) ? h->ldA : n;
    g.strideA = (what == 0) ? h->ldA * n : 3 * n * n;
    g.B = Zq; g.ldb = n; g.strideB = n * zcols;
    g.batch = nb;
    g.dot_out = od; g.dot_stride = nz;
    g.herm_dot = 1;
    g.dotX = Zq; g.lddx = n; g.strideX = n * zcols;
    wsb.ensure((size_t)gemm_ws_bytes(g));
    g.ws = wsb.p;
    launch_gemm<double>(g, s);
    if (!on_device)
      PPH_HIP(hipMemcpyAsync(out + q0 * nz, od, nb * nz * sizeof(double), hipMemcpyDeviceToHost, s));
    PPH_HIP(hipStreamSynchronize(s));
  }
  return PPH_OK;
  PPH_API_CATCH(h)
}

int pph_heevx(int64_t n, int64_t nmat, void* A, const pph_opts* opts, double* w, void* Z,
              int64_t* m, int32_t* info, int32_t on_device) {
  std::string why;
  if (check_opts(opts, why)) return fail(nullptr, "pph_heevx: " + why, PPH_ERR_ARG);
  if (n < 1 || nmat < 0 || (nmat > 0 && (!A || !w || !m || !info)))
    return fail(nullptr, "pph_heevx: bad arguments", PPH_ERR_ARG);
  if (Z && opts->jobz != 'V') return fail(nullptr, "pph_heevx: Z given with jobz = 'N'", PPH_ERR_ARG);
  Handle h;
  PPH_API_TRY
  h.opts = *opts;
  h.device = opts->device;
  PPH_HIP(hipSetDevice(h.device));
  PPH_HIP(hipStreamCreateWithFlags(&h.stream, hipStreamNonBlocking));
  h.set_n(n); h.N = 0;
  hipStream_t s = h.stream;
  const bool fp32 = opts->precision == 'S';
  const hipMemcpyKind in_kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  const hipMemcpyKind out_kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  const int64_t bmax = std::max<int64_t>(1, std::min<int64_t>(nmat, opts->batch > 0 ? opts->batch : 8));
  for (int64_t m0 = 0; m0 < nmat; m0 += bmax) {
    const int64_t B = std::min(bmax, nmat - m0);
    const int64_t ldA = h.ldA;
    h.A64.ensure((size_t)B * ldA * n * sizeof(cplx<double>));
    // the caller's matrices (leading dimension n) into the padded layout
    PPH_HIP(hipMemcpy2DAsync(h.A64.p, ldA * 16, (char*)A + (size_t)m0 * n * n * 16, n * 16, n * 16,
                             B * n, in_kind, s));
    EigResult R;
    if (fp32) {
      h.A32.ensure((size_t)B * ldA * n * sizeof(cplx<float>));
      launch_convert<double, float>(h.A64.as<cplx<double>>(), h.A32.as<cplx<float>>(), B * ldA * n, s);
      eig_batch<float>(h, h, B, h.A32.as<cplx<float>>(), h.A64.as<cplx<double>>(), Z != nullptr,
                       opts->vl, opts->vu, R);
    } else {
      eig_batch<double>(h, h, B, h.A64.as<cplx<double>>(), nullptr, Z != nullptr, opts->vl, opts->vu, R);
    }
    if (R.info_dev) {                        // the flags the batch left on the device
      std::vector<int32_t> fl(B);
      PPH_HIP(hipMemcpyAsync(fl.data(), R.info_dev, B * sizeof(int32_t), hipMemcpyDeviceToHost, s));
      PPH_HIP(hipStreamSynchronize(s));
      for (int64_t b = 0; b < B; ++b) R.info[b] |= fl[b];
    }
    for (int64_t b = 0; b < B; ++b) {
      const int64_t mm = R.M[b];
      if (mm > 0) {
        const int64_t f0 = b * R.mcap + R.first[b];
        PPH_HIP(hipMemcpyAsync(w + (m0 + b) * n, R.lam + f0, mm * sizeof(double), out_kind, s));
        if (Z && R.Z64)
          PPH_HIP(hipMemcpyAsync((char*)Z + (size_t)(m0 + b) * n * n * 16, R.Z64 + (size_t)f0 * n,
                                 (size_t)mm * n * 16, out_kind, s));
      }
      int32_t inf = R.info[b];
      if (on_device) {
        PPH_HIP(hipMemcpyAsync(m + m0 + b, &R.M[b], sizeof(int64_t), hipMemcpyHostToDevice, s));
        PPH_HIP(hipMemcpyAsync(info + m0 + b, &inf, sizeof(int32_t), hipMemcpyHostToDevice, s));
        PPH_HIP(hipStreamSynchronize(s));
      } else {
        m[m0 + b] = mm;
        info[m0 + b] = inf;
      }
    }
    PPH_HIP(hipStreamSynchronize(s));
  }
  return PPH_OK;
  PPH_API_CATCH(nullptr)
}

// ---------------------------------------------------------------- density of states
namespace {
// a host or device input array on the device
struct DevIn {
  DevBuf buf;
  const double* p = nullptr;
  void from(const void* src, size_t bytes, bool on_device, hipStream_t s) {
    if (on_device || bytes == 0) { p = static_cast<const double*>(src); return; }
    buf.ensure(bytes);
    PPH_HIP(hipMemcpyAsync(buf.p, src, bytes, hipMemcpyHostToDevice, s));
    p = buf.as<double>();
  }
};
struct Stream {
  hipStream_t s = nullptr;
  Stream() { PPH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
  ~Stream() { if (s) (void)hipStreamDestroy(s); }
};
bool ascending(const double* E, int64_t nE) {
  for (int64_t i = 1; i < nE; ++i)
    if (!(E[i - 1] <= E[i])) return false;
  return true;
}
}  // namespace

int pph_dos_triangulation(int64_t nk, int64_t nbands, const double* freq, const double* func,
                          int64_t ntri, const int64_t* tri, int64_t nE, const double* E, double* dos,
                          double* nd, int32_t device, int32_t on_device) {
  if (nk < 0 || nbands < 0 || ntri < 0 || nE < 0)
    return fail(nullptr, "pph_dos_triangulation: negative size", PPH_ERR_ARG);
  if (nE > 0 && (!E || !dos || !nd)) return fail(nullptr, "pph_dos_triangulation: E, dos and nd are required", PPH_ERR_ARG);
  if (ntri > 0 && nbands > 0 && (!freq || !tri))
    return fail(nullptr, "pph_dos_triangulation: freq and tri are required", PPH_ERR_ARG);
  if (!on_device) {
    for (int64_t i = 0; i < 3 * ntri; ++i)
      if (tri[i] < 0 || tri[i] >= nk) return fail(nullptr, "pph_dos_triangulation: vertex index out of range", PPH_ERR_ARG);
    if (!ascending(E, nE)) return fail(nullptr, "pph_dos_triangulation: energies must be ascending", PPH_ERR_ARG);
  }
  PPH_API_TRY
  PPH_HIP(hipSetDevice(device));
  Stream st;
  const hipStream_t s = st.s;
  const size_t fb = (size_t)nk * nbands * sizeof(double);
  DevIn df, dg, dE;
  DevBuf dtri, dout, dstep;
  df.from(freq, fb, on_device, s);
  // no func: the reference's default weight function, 1 everywhere (k_lt
  // reads a null weight array as ones)
  if (func) dg.from(func, fb, on_device, s);
  dE.from(E, nE * sizeof(double), on_device, s);
  const int64_t* dt = tri;
  if (!on_device && ntri > 0) {
    dtri.ensure(3 * ntri * sizeof(int64_t));
    PPH_HIP(hipMemcpyAsync(dtri.p, tri, 3 * ntri * sizeof(int64_t), hipMemcpyHostToDevice, s));
    dt = dtri.as<int64_t>();
  }
  double* od = dos;
  double* on = nd;
  if (!on_device) {
    dout.ensure(std::max<size_t>(16 * nE, 16));
    od = dout.as<double>();
    on = od + nE;
  }
  dstep.ensure(std::max<size_t>(8 * nE, 8));
  launch_dos_lt(df.p, dg.p, nbands, dt, ntri, dE.p, nE, od, on, dstep.as<double>(), s);
  if (!on_device && nE > 0) {
    PPH_HIP(hipMemcpyAsync(dos, od, nE * sizeof(double), hipMemcpyDeviceToHost, s));
    PPH_HIP(hipMemcpyAsync(nd, on, nE * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  PPH_HIP(hipStreamSynchronize(s));
  return PPH_OK;
  PPH_API_CATCH(nullptr)
}

// ---------------------------------------------------------------- no-allocation variant
// The difference array of the triangulation (the energies' "above the
// triangle" steps) lives in a file-scope device array instead of a buffer the
// library allocates, so the entry below allocates nothing.  Every GPU holds
// its own copy; host code uses the address hipGetSymbolAddress returns for the
// current device, never the array's name.
constexpr int64_t kDosScratchE = int64_t(1) << 20;       // energies (8 MiB of doubles)
static __device__ __align__(256) double scratch[kDosScratchE];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// this device's copy of `scratch` (looked up once per device, with it current)
double* dos_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int prev = 0;
    PPH_HIP(hipGetDevice(&prev));
    PPH_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    (void)hipSetDevice(prev);
    if (e != hipSuccess || !p)
      throw pph::hip_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  return g_scratch_ptr[device];
}
}  // namespace

int pph_dos_triangulation_noalloc(int64_t nk, int64_t nbands, const double* freq,
                                  const double* func, int64_t ntri, const int64_t* tri, int64_t nE,
                                  const double* E, double* dos, double* nd, int32_t device,
                                  void* stream) {
  if (nk < 0 || nbands < 0 || ntri < 0 || nE < 0)
    return fail(nullptr, "pph_dos_triangulation_noalloc: negative size", PPH_ERR_ARG);
  if (nE > kDosScratchE)
    return fail(nullptr, "pph_dos_triangulation_noalloc: nE = " + std::to_string(nE) +
                             " exceeds the scratch capacity of " + std::to_string(kDosScratchE) +
                             " energies (use pph_dos_triangulation)", PPH_ERR_ARG);
  if (nE > 0 && (!E || !dos || !nd))
    return fail(nullptr, "pph_dos_triangulation_noalloc: E, dos and nd are required", PPH_ERR_ARG);
  if (ntri > 0 && nbands > 0 && (!freq || !tri))
    return fail(nullptr, "pph_dos_triangulation_noalloc: freq and tri are required", PPH_ERR_ARG);
  PPH_API_TRY
  double* step = dos_scratch(device);
  int prev = 0;
  PPH_HIP(hipGetDevice(&prev));
  PPH_HIP(hipSetDevice(device));
  try {
    launch_dos_lt(freq, func, nbands, tri, ntri, E, nE, dos, nd, step,
                  static_cast<hipStream_t>(stream));
  } catch (...) {
    (void)hipSetDevice(prev);
    throw;
  }
  PPH_HIP(hipSetDevice(prev));
  return PPH_OK;
  PPH_API_CATCH(nullptr)
}

int pph_dos_smearing(int32_t method, int64_t nq, int64_t nbands, const double* freq,
                     const double* weight, double width, int64_t natom, int64_t nE, const double* E,
                     double* dos, int32_t device, int32_t on_device) {
  if (method != 1 && method != 2) return fail(nullptr, "pph_dos_smearing: method must be 1 (Gaussian) or 2 (Lorentzian)", PPH_ERR_ARG);
  if (nq < 0 || nbands < 0 || nE < 0 || natom < 1 || !(width > 0.0))
    return fail(nullptr, "pph_dos_smearing: bad sizes, natom or width", PPH_ERR_ARG);
  if (nE > 0 && (!E || !dos)) return fail(nullptr, "pph_dos_smearing: E and dos are required", PPH_ERR_ARG);
  if (nq > 0 && nbands > 0 && (!freq || !weight))
    return fail(nullptr, "pph_dos_smearing: freq and weight are required", PPH_ERR_ARG);
  if (!on_device) {
    if (!ascending(E, nE)) return fail(nullptr, "pph_dos_smearing: energies must be ascending", PPH_ERR_ARG);
    for (int64_t q = 0; q < nq; ++q)
      if (!ascending(freq + q * nbands, nbands))
        return fail(nullptr, "pph_dos_smearing: each q-point's frequencies must be ascending", PPH_ERR_ARG);
  }
  PPH_API_TRY
  PPH_HIP(hipSetDevice(device));
  Stream st;
  const hipStream_t s = st.s;
  DevIn df, dw, dE;
  DevBuf dout;
  df.from(freq, (size_t)nq * nbands * sizeof(double), on_device, s);
  dw.from(weight, nq * sizeof(double), on_device, s);
  dE.from(E, nE * sizeof(double), on_device, s);
  double* od = dos;
  if (!on_device) {
    dout.ensure(std::max<size_t>(8 * nE, 8));
    od = dout.as<double>();
  }
  launch_dos_smear(method, df.p, dw.p, nq, nbands, width, natom, dE.p, nE, od, s);
  if (!on_device && nE > 0) PPH_HIP(hipMemcpyAsync(dos, od, nE * sizeof(double), hipMemcpyDeviceToHost, s));
  PPH_HIP(hipStreamSynchronize(s));
  return PPH_OK;
  PPH_API_CATCH(nullptr)
}

int pph_fc_symmetrize(int64_t natom, double* fc, int64_t iterations, int32_t device,
                      int32_t on_device) {
  if (natom < 1 || iterations < 0 || !fc)
    return fail(nullptr, "pph_fc_symmetrize: bad arguments", PPH_ERR_ARG);
  PPH_API_TRY
  PPH_HIP(hipSetDevice(device));
  Stream st;
  const size_t bytes = (size_t)natom * natom * 9 * sizeof(double);
  DevBuf buf;
  double* d = fc;
  if (!on_device) {
    buf.ensure(bytes);
    PPH_HIP(hipMemcpyAsync(buf.p, fc, bytes, hipMemcpyHostToDevice, st.s));
    d = buf.as<double>();
  }
  launch_fc_symmetrize(d, natom, iterations, st.s);
  if (!on_device) PPH_HIP(hipMemcpyAsync(fc, d, bytes, hipMemcpyDeviceToHost, st.s));
  PPH_HIP(hipStreamSynchronize(st.s));
  return PPH_OK;
  PPH_API_CATCH(nullptr)
}

int pph_spring_fc(int64_t natom, int64_t npairs, const int64_t* pi, const int64_t* pj,
                  const double* rvec, const int32_t* kind, const double* params, uint64_t seed,
                  double noise_amp, double inter_r0, double inter_decay, double* fc,
                  int32_t device, int32_t on_device) {
  if (natom < 1 || npairs < 0 || !fc || !params ||
      (npairs > 0 && (!pi || !pj || !rvec || !kind)) || !(inter_decay > 0.0))
    return fail(nullptr, "pph_spring_fc: bad arguments", PPH_ERR_ARG);
  if (!on_device) {
    // a pair outside the cell or a kind the model does not know would write
    // outside fc or read outside the parameters
    for (int64_t p = 0; p < npairs; ++p)
      if (pi[p] < 0 || pi[p] >= natom || pj[p] < 0 || pj[p] >= natom || pi[p] == pj[p] ||
          kind[p] < 0 || kind[p] > 3)
        return fail(nullptr, "pph_spring_fc: pair or kind out of range", PPH_ERR_ARG);
  }
  PPH_API_TRY
  PPH_HIP(hipSetDevice(device));
  Stream st;
  SpringParams sp{};
  for (int k = 0; k < 7; ++k) sp.p[k] = params[k];
  sp.amp = noise_amp; sp.inter_r0 = inter_r0; sp.inter_decay = inter_decay; sp.seed = seed;
  const size_t bytes = (size_t)natom * natom * 9 * sizeof(double);
  DevBuf bfc, bi, bj, br, bk;
  double* d = fc;
  const int64_t *di = pi, *dj = pj;
  const double* dr = rvec;
  const int32_t* dk = kind;
  if (!on_device) {
    bfc.ensure(bytes);
    d = bfc.as<double>();
    if (npairs > 0) {
      bi.ensure(npairs * sizeof(int64_t)); bj.ensure(npairs * sizeof(int64_t));
      br.ensure(npairs * 3 * sizeof(double)); bk.ensure(npairs * sizeof(int32_t));
      PPH_HIP(hipMemcpyAsync(bi.p, pi, npairs * sizeof(int64_t), hipMemcpyHostToDevice, st.s));
      PPH_HIP(hipMemcpyAsync(bj.p, pj, npairs * sizeof(int64_t), hipMemcpyHostToDevice, st.s));
      PPH_HIP(hipMemcpyAsync(br.p, rvec, npairs * 3 * sizeof(double), hipMemcpyHostToDevice, st.s));
      PPH_HIP(hipMemcpyAsync(bk.p, kind, npairs * sizeof(int32_t), hipMemcpyHostToDevice, st.s));
      di = bi.as<int64_t>(); dj = bj.as<int64_t>(); dr = br.as<double>(); dk = bk.as<int32_t>();
    }
  }
  launch_spring_fc(natom, npairs, di, dj, dr, dk, sp, d, st.s);
  if (!on_device) PPH_HIP(hipMemcpyAsync(fc, d, bytes, hipMemcpyDeviceToHost, st.s));
  PPH_HIP(hipStreamSynchronize(st.s));
  return PPH_OK;
  PPH_API_CATCH(nullptr)
}

int pph_destroy(pph_handle* hh) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h) return PPH_OK;
  (void)hipSetDevice(h->device);
  for (size_t l = 0; l < h->lanes(); ++l)
    if (h->lane(l).stream) (void)hipStreamSynchronize(h->lane(l).stream);
  delete h;
  return PPH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- measurement hook
extern "C" int pph_debug_reduction(pph_handle* hh, int32_t mode, int32_t band) {
  if (mode < -1 || mode > 2)
    return fail(nullptr, "pph_debug_reduction: mode must be -1, 0, 1 or 2", PPH_ERR_ARG);
  if (band != 0 && band != 16 && band != 32)
    return fail(nullptr, "pph_debug_reduction: band must be 0 (keep), 16 or 32", PPH_ERR_ARG);
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h) {
    g_reduction_default = mode;
    if (band) g_band_default = band;
    return PPH_OK;
  }
  h->reduction = mode;
  if (band) h->band = band;
  return PPH_OK;
}

extern "C" int pph_debug_graphs(void* hh, int32_t mode, int64_t* captures, int64_t* replays) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h) return fail(nullptr, "pph_debug_graphs: handle is NULL", PPH_ERR_ARG);
  if (mode < -1 || mode > 1) return fail(h, "pph_debug_graphs: mode must be -1, 0 or 1", PPH_ERR_ARG);
  if (mode >= 0) h->graphs = mode == 1;
  int64_t c = 0, r = 0;
  for (size_t l = 0; l < h->lanes(); ++l) {
    c += h->lane(l).graph_captures;
    r += h->lane(l).graph_replays;
  }
  if (captures) *captures = c;
  if (replays) *replays = r;
  return PPH_OK;
}

extern "C" int pph_debug_gemm_tile(void* hh, int32_t tile) {
  if (tile < 0 || tile >= kGemmTileCount)
    return fail(nullptr, "pph_debug_gemm_tile: tile must be 0 (default), 1 (64x64), 2 (128x64), "
                         "3 (64x128) or 4 (128x128)", PPH_ERR_ARG);
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h) {
    set_gemm_tile_default(tile ? tile : (int)kGemmTile64x64);
    return PPH_OK;
  }
  if (tile == kGemmTile128x128 && h->opts.precision == 'D')
    return fail(h, "pph_debug_gemm_tile: the 128x128 tile is FP32 only", PPH_ERR_ARG);
  h->gemm_tile = tile;
  return PPH_OK;
}

extern "C" int pph_debug_cluster_rr(void* hh, int32_t mode) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (mode < 0 || mode > 1) return fail(h, "pph_debug_cluster_rr: mode must be 0 or 1", PPH_ERR_ARG);
  if (!h) {
    g_cluster_rr_default.store(mode);
    return PPH_OK;
  }
  h->cluster_rr = mode == 1;
  return PPH_OK;
}

extern "C" int pph_debug_vector_refinement(void* hh, int32_t steps, int32_t bf16) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (steps < 0 || steps > 2)
    return fail(h, "pph_debug_vector_refinement: steps must be 0, 1 or 2", PPH_ERR_ARG);
  if (bf16 < 0 || bf16 > 1)
    return fail(h, "pph_debug_vector_refinement: bf16 must be 0 or 1", PPH_ERR_ARG);
  if (!h) {
    g_vec_refine_default.store(steps);
    g_vec_bf16_default.store(bf16);
    return PPH_OK;
  }
  h->vec_refine = steps;
  h->vec_bf16 = bf16 == 1;
  return PPH_OK;
}

extern "C" int pph_debug_sparse(pph_handle* hh, int32_t mode, int64_t* nblk) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h) return fail(nullptr, "pph_debug_sparse: handle is NULL", PPH_ERR_ARG);
  if (mode < 0 || mode > 2) return fail(h, "pph_debug_sparse: mode must be 0, 1 or 2", PPH_ERR_ARG);
  if (mode == 2 && h->sp_nblk == 0)
    return fail(h, "pph_debug_sparse: no block lists (more than a quarter of the blocks nonzero)",
                PPH_ERR_ARG);
  h->sp_mode = mode;
  if (nblk) *nblk = h->use_sparse() ? h->sp_nblk : 0;
  return PPH_OK;
}

extern "C" int pph_debug_blocking(pph_handle* hh, int32_t nb_panel, int32_t nb_back) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h) return fail(nullptr, "pph_debug_blocking: handle is NULL", PPH_ERR_ARG);
  if (nb_panel < 16 || nb_panel > 64 || nb_panel % 16 != 0)
    return fail(h, "pph_debug_blocking: nb_panel must be 16, 32, 48 or 64", PPH_ERR_ARG);
  if (nb_back < 16 || nb_back > kVpad || nb_back % 16 != 0)
    return fail(h, "pph_debug_blocking: nb_back must be a multiple of 16 in [16, 128]", PPH_ERR_ARG);
  h->nb_panel = nb_panel;
  h->nb_back = nb_back;
  return PPH_OK;
}

extern "C" int pph_profile(pph_handle* hh, int32_t enable, int32_t reset, double* ms,
                           int64_t* counts) {
  Handle* h = reinterpret_cast<Handle*>(hh);
  if (!h) return fail(nullptr, "pph_profile: handle is NULL", PPH_ERR_ARG);
  PPH_API_TRY
  PPH_HIP(hipSetDevice(h->device));
  // summed over the lanes (phases of concurrent batches overlap in time)
  if (ms)
    for (int p = 0; p < PH_N; ++p) ms[p] = 0.0;
  if (counts)
    for (int p = 0; p < PH_N; ++p) counts[p] = 0;
  for (size_t l = 0; l < h->lanes(); ++l) {
    Timers& t = h->lane(l).tm;
    t.init();
    for (int p = 0; p < PH_N; ++p) {
      if (ms) ms[p] += t.ms[p];
      if (counts) counts[p] += t.cnt[p];
      if (reset) { t.ms[p] = 0; t.cnt[p] = 0; }
    }
    t.on = enable != 0;
  }
  return PPH_OK;
  PPH_API_CATCH(h)
}

// ---------------------------------------------------------------- test hook
extern "C" int pph_debug_zgemm(char prec, char opA, char opB, int64_t m, int64_t n, int64_t k,
                               const void* A, int64_t lda, const void* B, int64_t ldb, void* C,
                               int64_t ldc, int32_t mode, double* dot) {
  PPH_API_TRY
  const size_t es = (prec == 'S') ? 8 : 16;
  const int64_t ar = (opA == 'N') ? m : k, ac = (opA == 'N') ? k : m;
  const int64_t br = (opB == 'N') ? k : n, bc = (opB == 'N') ? n : k;
  (void)ar; (void)br;
  hipStream_t s;
  PPH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  DevBuf dA, dB, dC, dX, dW;
  dA.ensure(es * lda * ac); dB.ensure(es * ldb * bc); dC.ensure(es * ldc * n); dX.ensure(8 * n + 8);
  PPH_HIP(hipMemcpyAsync(dA.p, A, es * lda * ac, hipMemcpyHostToDevice, s));
  PPH_HIP(hipMemcpyAsync(dB.p, B, es * ldb * bc, hipMemcpyHostToDevice, s));
  PPH_HIP(hipMemcpyAsync(dC.p, C, es * ldc * n, hipMemcpyHostToDevice, s));
  launch_fill_zero(dX.p, 8 * n, s);
  auto run = [&](auto tag) {
    using T = decltype(tag);
    GemmParams<T> g{};
    g.opA = opA; g.opB = opB; g.m = m; g.n = n; g.k = k;
    g.alpha = cplx<T>{1, 0}; g.beta = cplx<T>{0, 0};
    g.A = dA.as<cplx<T>>(); g.lda = lda; g.B = dB.as<cplx<T>>(); g.ldb = ldb;
    g.C = dC.as<cplx<T>>(); g.ldc = ldc; g.batch = 1;
    if (mode == 1 || mode == 4) {
      // mode 4: A Hermitian (lower triangle read), herm_dot epilogue
      g.dot_out = dX.as<double>(); g.dotX = dC.as<cplx<T>>(); g.lddx = ldc;
      g.herm_dot = (mode == 4) ? 1 : 0;
    } else if (mode == 2) {
      g.ksplit = 4;
    } else if (mode == 3) {
      // C = A(:, 0:k/2) B(0:k/2, :) + A(:, k/2:) B(k/2:, :) through the two-product path (opA = opB = 'N')
      g.k1 = k / 2; g.A2 = dA.as<cplx<T>>() + (k / 2) * lda; g.B2 = dB.as<cplx<T>>() + k / 2;
      g.strideA2 = 0; g.strideB2 = 0;
    } else if (mode == 5) {
      // C = A B for Hermitian A (m = k, opA = 'N') from its lower triangle:
      // the lower triangle with the diagonal, then its strict part transposed
      g.tri = 1;
      g.four_m = 1;
      launch_gemm<T>(g, s);
      g.opA = 'C'; g.tri = 2; g.beta = cplx<T>{1, 0};
    }
    dW.ensure((size_t)gemm_ws_bytes(g));
    g.ws = dW.p;
    if constexpr (std::is_same_v<T, float>) {
      if (mode == 6) {                   // the BF16 matrix-core path (complex64 in and out)
        launch_cgemm_bf16(g, s);
        return;
      }
    }
    launch_gemm<T>(g, s);
  };
  if (mode == 6 && prec != 'S') return fail(nullptr, "pph_debug_zgemm: mode 6 is complex64 only", PPH_ERR_ARG);
  if (prec == 'S') run(float{}); else run(double{});
  PPH_HIP(hipMemcpyAsync(C, dC.p, es * ldc * n, hipMemcpyDeviceToHost, s));
  if (dot) PPH_HIP(hipMemcpyAsync(dot, dX.p, 8 * n, hipMemcpyDeviceToHost, s));
  PPH_HIP(hipStreamSynchronize(s));
  PPH_HIP(hipStreamDestroy(s));
  return PPH_OK;
  PPH_API_CATCH(nullptr)
}
