// Driver of the tiled Cholesky graph on the MI355X runtime, with the
// command line and output of examples/potrf/testing_dpotrf.cc:
//   -N <n> (default 160)  -t <nb> (default 32)  -n <niter> (default 3)
//   -c <threads>  -x (skip the check)  -w (no defer-writer hint)
//   -P/-Q/-dag/--default-device-map are accepted and ignored (one process, one GPU).
// Without -x: PLGSY(bump=N, seed=3872), factor, and the check
// ||L L^T - A||_inf / (||A||_inf N eps) <= 60 of testing_dpotrf.cc:277-335.
// With -x: niter timed runs, each "TTG Execution Time (milliseconds) : t :
// Flops F G GF/s" (generation, factorization and write-back timed, as in the
// reference).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "plgsy.h"
#include "pmw.h"
#include "potrf.h"
#include "result.h"
#include "ttg.h"

static const char *get_opt(int argc, char **argv, const char *name) {
  for (int i = 1; i + 1 < argc; ++i)
    if (!std::strcmp(argv[i], name)) return argv[i + 1];
  return nullptr;
}
static bool has_opt(int argc, char **argv, const char *name) {
  for (int i = 1; i < argc; ++i)
    if (!std::strcmp(argv[i], name)) return true;
  return false;
}

// ||L L^T - A0||_inf / (||A0||_inf N eps): L in the lower part of A (dense,
// column-major; its upper part is not read), A0 the original matrix as
// getLAPACKMatrix returns it; the residual is mirrored from its lower part
// and ||A0||_inf is taken of A0 as given (testing_dpotrf.cc:277-335).
static int check_dpotrf(const std::vector<double> &A, const std::vector<double> &A0, int N) {
  std::vector<double> R((size_t)N * N);
  for (int j = 0; j < N; ++j)
    for (int i = j; i < N; ++i) {
      double s = 0.0;
      for (int k = 0; k <= j; ++k) s += A[i + (size_t)k * N] * A[j + (size_t)k * N];
      R[i + (size_t)j * N] = s - A0[i + (size_t)j * N];
    }
  double anorm = 0.0, rnorm = 0.0;
  for (int i = 0; i < N; ++i) {
    double sa = 0.0, sr = 0.0;
    for (int j = 0; j < N; ++j) {
      sa += std::fabs(A0[i + (size_t)j * N]);
      sr += std::fabs(i >= j ? R[i + (size_t)j * N] : R[j + (size_t)i * N]);
    }
    anorm = std::max(anorm, sa);
    rnorm = (sr > rnorm || std::isnan(sr)) ? sr : rnorm;
  }
  const double result = rnorm / (anorm * N * std::numeric_limits<double>::epsilon());
  std::cout << "============" << std::endl;
  std::cout << "Checking the Cholesky factorization" << std::endl;
  std::cout << "-- ||A||_oo = " << anorm << ", ||LL'-A||_oo = " << rnorm << std::endl;
  std::cout << "-- ||LL'-A||_oo/(||A||_oo.N.eps) = " << result << std::endl;
  if (std::isnan(rnorm) || std::isinf(rnorm) || std::isnan(result) || std::isinf(result) || result > 60.0) {
    std::cout << "-- Factorization is suspicious !" << std::endl;
    return -1;
  }
  std::cout << "-- Factorization is CORRECT !" << std::endl;
  return 0;
}

int main(int argc, char **argv) {
  int NB = 32, N = 5 * NB, nthreads = -1, niter = 3;
  if (const char *o = get_opt(argc, argv, "-N")) N = std::atoi(o);
  if (const char *o = get_opt(argc, argv, "-t")) NB = std::atoi(o);
  if (const char *o = get_opt(argc, argv, "-c")) nthreads = std::atoi(o);
  if (const char *o = get_opt(argc, argv, "-n")) niter = std::atoi(o);
  const bool check = !has_opt(argc, argv, "-x");
  const bool cow_hint = !has_opt(argc, argv, "-w");
  const unsigned long seed = 3872;

  ttg::initialize(1, argv, nthreads);
  auto world = ttg::default_execution_context();
  std::cout << "Creating 2D block cyclic matrix with NB " << NB << " N " << N << " M " << N << " P 1 Q 1"
            << std::endl;
  MatrixT<double> A(N, NB);
  int ret = EXIT_SUCCESS;

  auto startup_tt = [&](ttg::Edge<Key2, void> &startup) {
    auto init_tt = ttg::make_tt<void>(
        [&](std::tuple<ttg::Out<Key2, void>> &out) {
          for (int i = 0; i < (int)A.rows(); i++)
            for (int j = 0; j <= i && j < (int)A.cols(); j++) ttg::sendk<0>(Key2{i, j}, out);
        },
        ttg::edges(), ttg::edges(startup), "Startup Trigger", {}, {"startup"});
    return init_tt;
  };

  if (!check) {
    for (int it = 0; it < niter; ++it) {
      ttg::Edge<Key2, void> startup("startup");
      ttg::Edge<Key2, MatrixTile<double>> topotrf("To POTRF"), result("To result");
      auto init_tt = startup_tt(startup);
      auto plgsy_ttg = make_plgsy_ttg(A, N, seed, startup, topotrf, cow_hint);
      auto potrf_ttg = potrf::make_potrf_ttg(A, topotrf, result, cow_hint);
      auto result_ttg = make_result_ttg(A, result, cow_hint);
      ttg::make_graph_executable(init_tt.get());
      auto beg = std::chrono::high_resolution_clock::now();
      init_tt->invoke();
      ttg::execute(world);
      ttg::fence(world);
      auto end = std::chrono::high_resolution_clock::now();
      if (ttg::cancelled()) break;
      const double us = std::chrono::duration_cast<std::chrono::microseconds>(end - beg).count();
      std::cout << "TTG Execution Time (milliseconds) : " << us / 1e3 << " : Flops " << potrf::FLOPS_DPOTRF(N) << " "
                << (potrf::FLOPS_DPOTRF(N) / 1e9) / (us / 1e6) << " GF/s" << std::endl;
    }
  } else {
    {
      ttg::Edge<Key2, void> startup("startup");
      ttg::Edge<Key2, MatrixTile<double>> result("To result");
      auto init_tt = startup_tt(startup);
      auto plgsy_ttg = make_plgsy_ttg(A, N, seed, startup, result, cow_hint);
      auto result_ttg = make_result_ttg(A, result, cow_hint);
      ttg::make_graph_executable(init_tt.get());
      init_tt->invoke();
      ttg::execute(world);
      ttg::fence(world);
    }
    std::vector<double> A0 = A.getLAPACKMatrix();
    {
      ttg::Edge<Key2, MatrixTile<double>> topotrf("To POTRF"), toresult("To Result");
      auto load = make_load_tt(A, topotrf, cow_hint);
      auto potrf_ttg = potrf::make_potrf_ttg(A, topotrf, toresult, cow_hint);
      auto result_ttg = make_result_ttg(A, toresult, cow_hint);
      ttg::make_graph_executable(load.get());
      load->invoke();
      ttg::fence(world);
    }
    if (!ttg::cancelled()) {
      std::vector<double> L = A.getLAPACKMatrix();
      if (check_dpotrf(L, A0, N) == -1) ret = EXIT_FAILURE;
    }
  }
  if (ttg::cancelled()) {
    std::cout << "Factorization stopped: info = " << ttg::cancel_info() << std::endl;
    ret = EXIT_FAILURE;
  }
  auto &rt = ttg::detail::Runtime::get();
  std::cout << "runtime: tasks " << rt.tasks_run << " rounds " << rt.rounds << " kernel launches "
            << rt.kernel_launches << " largest round " << rt.max_round << std::endl;
  ttg::finalize();
  return ret;
}
