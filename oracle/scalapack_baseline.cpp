// CPU baseline (test infrastructure, not part of the library): the per-q
// eigensolve and velocity products of parphom.x run with MKL ScaLAPACK on
// host cores, one MPI rank per core, with the reference's argument choices:
//   pzheevx('V', 'A', 'U', 3N, D, ...) with a workspace query first
//   (Phonon_Calculations/SRC/diagonalize.F90:73-96), abstol and orfac as
//   given, block size mb = nb (the examples' inp.dat use 32);
//   per Cartesian direction, pzgemm('N','N') dD E and pzgemm('C','N') E^H (dD E)
//   (write_output.F90:415-421; the velocity is the diagonal of the second);
//   the process grid of get_optimum_grid: nprow x npcol = P with nprow the
//   largest divisor of P not above sqrt(P), column-major (blacs.F90:203-248).
// Input: a file with, per q-point, D(q) and dD/dq_{x,y,z}, each 3N x 3N
// complex128 column-major (written by bench.py's cpu_baseline leg from the
// oracle's assembly).  Output on rank 0: one line of JSON with the seconds
// per q-point of the eigensolve and of the velocity products, and the
// lowest eigenvalue of each q (a check against the oracle).
//
// Build: see oracle/Makefile (target scalapack_baseline, mpicxx-free: g++ with
// the MPICH and MKL libraries of /opt/conda).
#include <mpi.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

typedef std::complex<double> zc;

extern "C" {
void blacs_get_(const int* ctxt, const int* what, int* val);
void blacs_gridinit_(int* ctxt, const char* order, const int* nprow, const int* npcol);
void blacs_gridinfo_(const int* ctxt, int* nprow, int* npcol, int* myrow, int* mycol);
void blacs_gridexit_(const int* ctxt);
int numroc_(const int* n, const int* nb, const int* iproc, const int* isrcproc, const int* nprocs);
void descinit_(int* desc, const int* m, const int* n, const int* mb, const int* nb, const int* irsrc,
               const int* icsrc, const int* ictxt, const int* lld, int* info);
void pzheevx_(const char* jobz, const char* range, const char* uplo, const int* n, zc* a,
              const int* ia, const int* ja, const int* desca, const double* vl, const double* vu,
              const int* il, const int* iu, const double* abstol, int* m, int* nz, double* w,
              const double* orfac, zc* z, const int* iz, const int* jz, const int* descz, zc* work,
              const int* lwork, double* rwork, const int* lrwork, int* iwork, const int* liwork,
              int* ifail, int* iclustr, double* gap, int* info);
void pzgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k,
             const zc* alpha, const zc* a, const int* ia, const int* ja, const int* desca,
             const zc* b, const int* ib, const int* jb, const int* descb, const zc* beta, zc* c,
             const int* ic, const int* jc, const int* descc);
}

// global index of local index l (0-based) on process p of np, block nb
static long l2g(long l, int nb, int p, int np) { return (l / nb) * (long)nb * np + (long)p * nb + l % nb; }

int main(int argc, char** argv) {
  MPI_Init(&argc, &argv);
  int rank = 0, P = 1;
  MPI_Comm_rank(MPI_COMM_WORLD, &rank);
  MPI_Comm_size(MPI_COMM_WORLD, &P);
  if (argc < 4) {
    if (rank == 0) std::fprintf(stderr, "usage: %s file n nq [nb abstol orfac]\n", argv[0]);
    MPI_Finalize();
    return 2;
  }
  const char* path = argv[1];
  const int n = std::atoi(argv[2]), nq = std::atoi(argv[3]);
  const int nb = argc > 4 ? std::atoi(argv[4]) : 32;
  const double abstol = argc > 5 ? std::atof(argv[5]) : (double)1e-15f;
  const double orfac = argc > 6 ? std::atof(argv[6]) : (double)1e-15f;
  // the most square grid with nprow <= npcol
  int nprow = 1;
  for (int d = 1; d * d <= P; ++d)
    if (P % d == 0) nprow = d;
  int npcol = P / nprow;
  int ctxt = 0, zero = 0, one = 1, myrow = 0, mycol = 0, info = 0;
  blacs_get_(&zero, &zero, &ctxt);
  blacs_gridinit_(&ctxt, "C", &nprow, &npcol);
  blacs_gridinfo_(&ctxt, &nprow, &npcol, &myrow, &mycol);
  const int locr = numroc_(&n, &nb, &myrow, &zero, &nprow);
  const int locc = numroc_(&n, &nb, &mycol, &zero, &npcol);
  const int lld = std::max(1, locr);
  int desc[9];
  descinit_(desc, &n, &n, &nb, &nb, &zero, &zero, &ctxt, &lld, &info);
  const size_t loc = (size_t)lld * locc;
  std::vector<zc> A(loc), Z(loc), Y(loc), Vm(loc), dD(3 * loc);
  std::vector<double> w(n);
  std::vector<zc> col(n);
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    std::fprintf(stderr, "rank %d: cannot open %s\n", rank, path);
    MPI_Abort(MPI_COMM_WORLD, 3);
  }
  // this rank's entries of matrix k of q-point iq: its columns read whole
  auto read_local = [&](int iq, int k, zc* dst) {
    const long base = ((long)iq * 4 + k) * (long)n * n;
    for (int jl = 0; jl < locc; ++jl) {
      const long j = l2g(jl, nb, mycol, npcol);
      std::fseek(f, (base + j * (long)n) * (long)sizeof(zc), SEEK_SET);
      if (std::fread(col.data(), sizeof(zc), n, f) != (size_t)n) MPI_Abort(MPI_COMM_WORLD, 4);
      for (int il = 0; il < locr; ++il) dst[il + (size_t)jl * lld] = col[l2g(il, nb, myrow, nprow)];
    }
  };
  double t_eig = 0.0, t_vel = 0.0;
  std::vector<double> lowest;
  for (int iq = 0; iq < nq; ++iq) {
    read_local(iq, 0, A.data());
    for (int a = 0; a < 3; ++a) read_local(iq, 1 + a, dD.data() + a * loc);
    MPI_Barrier(MPI_COMM_WORLD);
    auto t0 = std::chrono::steady_clock::now();
    // pzheevx with a workspace query, as diagonalize.F90 calls it
    int m = 0, nz = 0;
    double vl = 0.0, vu = 0.0;
    std::vector<int> ifail(n), iclustr(2 * nprow * npcol);
    std::vector<double> gap(nprow * npcol);
    int lwork = -1, lrwork = -1, liwork = -1;
    zc wq;
    double rq;
    int iq_ = 0;
    pzheevx_("V", "A", "U", &n, A.data(), &one, &one, desc, &vl, &vu, &one, &one, &abstol, &m, &nz,
             w.data(), &orfac, Z.data(), &one, &one, desc, &wq, &lwork, &rq, &lrwork, &iq_, &liwork,
             ifail.data(), iclustr.data(), gap.data(), &info);
    lwork = (int)std::abs(wq.real()) + 1;
    lrwork = (int)std::abs(rq) + 1;
    liwork = std::abs(iq_) + 1;
    std::vector<zc> work(lwork);
    std::vector<double> rwork(lrwork);
    std::vector<int> iwork(liwork);
    pzheevx_("V", "A", "U", &n, A.data(), &one, &one, desc, &vl, &vu, &one, &one, &abstol, &m, &nz,
             w.data(), &orfac, Z.data(), &one, &one, desc, work.data(), &lwork, rwork.data(), &lrwork,
             iwork.data(), &liwork, ifail.data(), iclustr.data(), gap.data(), &info);
    MPI_Barrier(MPI_COMM_WORLD);
    auto t1 = std::chrono::steady_clock::now();
    // velocity products: dD E, then E^H (dD E), per direction
    const zc alpha(1.0, 0.0), beta(0.0, 0.0);
    for (int a = 0; a < 3; ++a) {
      pzgemm_("N", "N", &n, &n, &n, &alpha, dD.data() + a * loc, &one, &one, desc, Z.data(), &one, &one,
              desc, &beta, Y.data(), &one, &one, desc);
      pzgemm_("C", "N", &n, &n, &n, &alpha, Z.data(), &one, &one, desc, Y.data(), &one, &one, desc,
              &beta, Vm.data(), &one, &one, desc);
    }
    MPI_Barrier(MPI_COMM_WORLD);
    auto t2 = std::chrono::steady_clock::now();
    t_eig += std::chrono::duration<double>(t1 - t0).count();
    t_vel += std::chrono::duration<double>(t2 - t1).count();
    lowest.push_back(info == 0 ? w[0] : std::nan(""));
    if (info != 0 && rank == 0) std::fprintf(stderr, "pzheevx info %d at q %d\n", info, iq);
  }
  std::fclose(f);
  if (rank == 0) {
    std::printf("{\"ranks\": %d, \"nprow\": %d, \"npcol\": %d, \"nb\": %d, \"n\": %d, \"nq\": %d, "
                "\"eig_s_per_q\": %.6f, \"vel_s_per_q\": %.6f, \"lowest_lambda\": [",
                P, nprow, npcol, nb, n, nq, t_eig / nq, t_vel / nq);
    for (size_t i = 0; i < lowest.size(); ++i) std::printf("%s%.17g", i ? ", " : "", lowest[i]);
    std::printf("]}\n");
  }
  blacs_gridexit_(&ctxt);
  MPI_Finalize();
  return 0;
}
