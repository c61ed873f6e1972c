// POINV driver (POTRF then POTRI, or POTRF then TRTRI and LAUUM) on the
// device drivers of libttg_chol, with the command line and output lines of
// examples/potrf/testing_dpoinv.cc:
//   -N <n> (default 160)  -t <nb> (default 32)  -nruns <r> (default 3)
//   -seq 0: POTRF + POTRI; -seq 1: POTRF + TRTRI + LAUUM, each timed; -seq 2:
//   the whole POINV timed as one (the reference's unsplit graph)
//   -x: no check.  -P/-Q/-dag/-v/-ttg-dag are accepted and ignored (one
//   process, one GPU).
//   -graph: the reference's task graphs instead (potrf.h, potri.h, trtri_L.h,
//   lauum.h) on the TTG runtime, over a host matrix, each step timed from
//   its load to its fence as in testing_dpoinv.cc:185-330; -c sets the
//   runtime's threads and -w drops the defer-writer hint.  Without it, the
//   planned device drivers of libttg_chol run (tiles stay in device memory).
// Each run generates PLGSY(bump = N, seed = 3872) in the lower tiles; run 0
// is a warm-up, as in the reference (results are printed for runs 1..nruns).
// The check (N <= 4096): ||I - A A^{-1}||_1 / (||A||_1 ||A^{-1}||_1 N eps) <= 10,
// the measure of check_dtrtri (testing_dtrtri.cc:269-325) applied to the
// inverse of the symmetric matrix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "lauum.h"
#include "plgsy.h"
#include "pmw.h"
#include "potrf.h"
#include "potri.h"
#include "result.h"
#include "trtri_L.h"
#include "ttg.h"
#include "ttg_chol.h"

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

// DPLASMA flop counts (potrf.h:30-33, lauum.h:11-25)
static double flops_dpotrf(double n) { return n * n * n / 3.0 + n * n / 2.0 + n / 6.0; }
static double flops_dtrtri(double n) {
  return n * (n * (1. / 6. * n + 0.5) + 1. / 3.) + n * (n * (1. / 6. * n - 0.5) + 1. / 3.);
}
static double flops_dpotri(double n) {
  return n * ((2. / 3.) + n * ((1. / 3.) * n + 1.)) + n * ((1. / 6.) + n * ((1. / 3.) * n - 0.5));
}

static long tile_index(int m, int n, int nt) { return (long)n * nt - (long)n * (n - 1) / 2 + (m - n); }

// dense symmetric N x N (column-major) from its lower tiles
static std::vector<double> dense_sym(const std::vector<double> &t, int N, int NB) {
  const int nt = (N + NB - 1) / NB;
  std::vector<double> A((size_t)N * N);
  for (int j = 0; j < N; ++j)
    for (int i = j; i < N; ++i) {
      const double v = t[tile_index(i / NB, j / NB, nt) * (size_t)NB * NB + (i % NB) + (size_t)(j % NB) * NB];
      A[i + (size_t)j * N] = v;
      A[j + (size_t)i * N] = v;
    }
  return A;
}

static double norm1(const std::vector<double> &A, int N) {
  double m = 0.0;
  for (int j = 0; j < N; ++j) {
    double s = 0.0;
    for (int i = 0; i < N; ++i) s += std::fabs(A[i + (size_t)j * N]);
    m = std::max(m, s);
  }
  return m;
}

static int check_dpoinv(const std::vector<double> &A0t, const std::vector<double> &Xt, int N, int NB) {
  const std::vector<double> A = dense_sym(A0t, N, NB), X = dense_sym(Xt, N, NB);
  std::vector<double> R((size_t)N * N, 0.0);
  for (int j = 0; j < N; ++j) {
    for (int k = 0; k < N; ++k) {
      const double x = X[k + (size_t)j * N];
      const double *a = &A[(size_t)k * N];
      double *r = &R[(size_t)j * N];
      for (int i = 0; i < N; ++i) r[i] -= a[i] * x;
    }
    R[j + (size_t)j * N] += 1.0;
  }
  const double an = norm1(A, N), xn = norm1(X, N), rn = norm1(R, N);
  const double result = (rn / an / xn) / (std::numeric_limits<double>::epsilon() * N);
  std::cout << "============" << std::endl;
  std::cout << "Checking POINV" << std::endl;
  std::cout << "-- ||A||_one = " << an << " ||A^(-1)||_one = " << xn << " ||I - A * A^(-1)||_one = " << rn
            << ", result = " << result << std::endl;
  if (std::isinf(xn) || std::isnan(result) || std::isinf(result) || result > 10.0) {
    std::cout << "-- Inversion is suspicious !" << std::endl;
    return -1;
  }
  std::cout << "-- Inversion is CORRECT !" << std::endl;
  return 0;
}

#define CHECK_RC(call)                                                                    \
  do {                                                                                    \
    int _rc = (call);                                                                     \
    if (_rc != TTG_CHOL_OK) {                                                             \
      std::cerr << #call << " failed: " << _rc << " " << ttg_chol_last_error() << std::endl; \
      return EXIT_FAILURE;                                                                \
    }                                                                                     \
  } while (0)

// The task-graph form: PLGSY by its graph (untimed), then POTRF and POTRI
// (seq 0), POTRF, TRTRI and LAUUM (seq 1) or POTRF and POTRI as one graph
// (seq 2), each a load -> graph -> result writer over the host matrix.
static int run_graphs(int N, int NB, int nruns, int seq, int nthreads, bool check, bool cow) {
  ttg::initialize(1, nullptr, nthreads);
  MatrixT<double> A(N, NB);
  const int nt = (N + NB - 1) / NB;
  const size_t nel = (size_t)nt * (nt + 1) / 2 * NB * NB;
  using clk = std::chrono::high_resolution_clock;
  auto ms = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration_cast<std::chrono::microseconds>(b - a).count() / 1e3;
  };
  const std::string head = std::string(" (") + (cow ? "with" : "without") + " defer writer) -- N= " +
                           std::to_string(N) + " NB= " + std::to_string(NB) + " P= 1 Q= 1 nthreads= " +
                           std::to_string(ttg::detail::Runtime::get().num_threads());
  auto line = [&](const char *kind, const char *what, double t_ms, double fl) {
    std::cout << "POINV (" << kind << ")" << head << " " << what << "TTG Execution Time (milliseconds) : " << t_ms
              << " : Flops " << fl << " " << (fl / 1e9) / (t_ms / 1e3) << " GF/s" << std::endl;
  };
  // one load -> graph -> result writer pass; false if the graph stopped
  auto step = [&](auto make) {
    ttg::Edge<Key2, MatrixTile<double>> in("To graph"), out("To result");
    auto load = make_load_tt(A, in, cow);
    auto g = make(in, out);
    auto res = make_result_ttg(A, out, cow);
    ttg::make_graph_executable(load.get());
    load->invoke();
    ttg::execute();
    ttg::fence();
    return !ttg::cancelled();
  };
  auto potrf_g = [&](auto &in, auto &out) { return potrf::make_potrf_ttg(A, in, out, cow); };
  auto potri_g = [&](auto &in, auto &out) { return potri::make_potri_ttg(A, in, out, cow); };
  auto trtri_g = [&](auto &in, auto &out) { return trtri_lower::make_trtri_ttg(A, in, out, cow); };
  auto lauum_g = [&](auto &in, auto &out) { return lauum::make_lauum_ttg(A, in, out, cow); };
  auto poinv_g = [&](auto &in, auto &out) {
    ttg::Edge<Key2, MatrixTile<double>> mid("POTRF to POTRI");
    auto f = potrf::make_potrf_ttg(A, in, mid, cow);
    auto i = potri::make_potri_ttg(A, mid, out, cow);
    std::vector<std::unique_ptr<ttg::TTBase>> ops;
    ops.push_back(std::move(f));
    ops.push_back(std::move(i));
    return ttg::make_ttg(std::move(ops), std::make_tuple(), std::make_tuple(), "POINV TTG");
  };
  int ret = EXIT_SUCCESS;
  std::vector<double> A0;
  for (int t = 0; t <= nruns && ret == EXIT_SUCCESS; ++t) {
    {
      ttg::Edge<Key2, void> startup("startup");
      ttg::Edge<Key2, MatrixTile<double>> res("To result");
      auto init = ttg::make_tt<void>(
          [&](std::tuple<ttg::Out<Key2, void>> &o) {
            for (int i = 0; i < nt; i++)
              for (int j = 0; j <= i; j++) ttg::sendk<0>(Key2{i, j}, o);
          },
          ttg::edges(), ttg::edges(startup), "Startup Trigger");
      auto plgsy = make_plgsy_ttg(A, (unsigned long)N, 3872, startup, res, cow);
      auto store = make_result_ttg(A, res, cow);
      init->invoke();
      ttg::fence();
    }
    if (check && t == nruns && N <= 4096) A0.assign(A.tile_ptr(0, 0), A.tile_ptr(0, 0) + nel);
    bool ok = true;
    const auto b0 = clk::now();
    if (seq == 2) {
      ok = step(poinv_g);
      const auto e = clk::now();
      if (ok && t > 0) line("POINV", "", ms(b0, e), flops_dpotrf(N) + flops_dpotri(N));
    } else {
      ok = step(potrf_g);
      const auto b1 = clk::now();
      if (ok && seq == 0) {
        ok = step(potri_g);
        const auto e = clk::now();
        if (ok && t > 0) {
          line("POTRF+POTRI", "POTRF ", ms(b0, b1), flops_dpotrf(N));
          line("POTRF+POTRI", "POTRI ", ms(b1, e), flops_dpotri(N));
          line("POTRF+POTRI", "", ms(b0, e), flops_dpotrf(N) + flops_dpotri(N));
        }
      } else if (ok) {
        ok = step(trtri_g);
        const auto b2 = clk::now();
        ok = ok && step(lauum_g);
        const auto e = clk::now();
        if (ok && t > 0) {
          line("POTRF+TRTRI+LAUUM", "POTRF ", ms(b0, b1), flops_dpotrf(N));
          line("POTRF+TRTRI+LAUUM", "TRTRI ", ms(b1, b2), flops_dtrtri(N));
          line("POTRF+TRTRI+LAUUM", "LAUUM ", ms(b2, e), flops_dpotri(N) - flops_dtrtri(N));
          line("POTRF+TRTRI+LAUUM", "", ms(b0, e), flops_dpotrf(N) + flops_dpotri(N));
        }
      }
    }
    if (!ok) {
      std::cout << "POINV failed: info = " << ttg::cancel_info() << std::endl;
      ret = EXIT_FAILURE;
    }
  }
  if (ret == EXIT_SUCCESS && !A0.empty()) {
    const std::vector<double> X(A.tile_ptr(0, 0), A.tile_ptr(0, 0) + nel);
    if (check_dpoinv(A0, X, N, NB) != 0) ret = EXIT_FAILURE;
  }
  auto &rt = ttg::detail::Runtime::get();
  std::cout << "runtime: tasks " << rt.tasks_run << " rounds " << rt.rounds << " kernel launches "
            << rt.kernel_launches << " largest round " << rt.max_round << std::endl;
  ttg::finalize();
  return ret;
}

int main(int argc, char **argv) {
  int NB = 32, N = 5 * NB, nruns = 3, seq = 0, nthreads = -1;
  if (const char *o = get_opt(argc, argv, "-N")) N = std::atoi(o);
  if (const char *o = get_opt(argc, argv, "-t")) NB = std::atoi(o);
  if (const char *o = get_opt(argc, argv, "-nruns")) nruns = std::atoi(o);
  if (const char *o = get_opt(argc, argv, "-seq")) seq = std::atoi(o);
  if (const char *o = get_opt(argc, argv, "-c")) nthreads = std::atoi(o);
  const bool check = !has_opt(argc, argv, "-x");
  const bool cow = !has_opt(argc, argv, "-w");
  const unsigned long long seed = 3872;
  const int nt = (N + NB - 1) / NB;
  const size_t nel = (size_t)nt * (nt + 1) / 2 * NB * NB;
  std::cout << "Creating 2D block cyclic matrix with NB " << NB << " N " << N << " M " << N << " P 1" << std::endl;
  if (has_opt(argc, argv, "-graph")) return run_graphs(N, NB, nruns, seq, nthreads, check, cow);

  double *tiles = nullptr;
  int *info = nullptr;
  if (hipMalloc(&tiles, nel * sizeof(double)) != hipSuccess || hipMalloc(&info, sizeof(int)) != hipSuccess) {
    std::cerr << "device allocation failed" << std::endl;
    return EXIT_FAILURE;
  }
  using clk = std::chrono::high_resolution_clock;
  auto ms = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration_cast<std::chrono::microseconds>(b - a).count() / 1e3;
  };
  auto sync = []() { return hipDeviceSynchronize() == hipSuccess ? 0 : -1; };
  const std::string head = std::string(" (") + (cow ? "with" : "without") + " defer writer) -- N= " +
                           std::to_string(N) + " NB= " + std::to_string(NB) + " P= 1 Q= 1 nthreads= " +
                           std::to_string(nthreads);
  auto line = [&](const char *kind, const char *what, double t_ms, double fl) {
    std::cout << "POINV (" << kind << ")" << head << " " << what << "TTG Execution Time (milliseconds) : " << t_ms
              << " : Flops " << fl << " " << (fl / 1e9) / (t_ms / 1e3) << " GF/s" << std::endl;
  };
  int ret = EXIT_SUCCESS;
  std::vector<double> A0;
  for (int t = 0; t <= nruns; ++t) {
    CHECK_RC(ttg_chol_plgsy_tiled(nullptr, N, NB, (double)N, seed, tiles));
    if (sync()) return EXIT_FAILURE;
    if (check && t == nruns && N <= 4096) {
      A0.resize(nel);
      if (hipMemcpy(A0.data(), tiles, nel * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return EXIT_FAILURE;
    }
    const auto b0 = clk::now();
    CHECK_RC(ttg_chol_dpotrf_tiled(nullptr, N, NB, tiles, info, TTG_CHOL_F_DEFAULT));
    if (seq == 2) {
      CHECK_RC(ttg_chol_dpotri_tiled(nullptr, N, NB, tiles, info));
      if (sync()) return EXIT_FAILURE;
      const auto e = clk::now();
      if (t > 0) line("POINV", "", ms(b0, e), flops_dpotrf(N) + flops_dpotri(N));
    } else {
      if (sync()) return EXIT_FAILURE;
      const auto b1 = clk::now();
      if (seq == 0) {
        CHECK_RC(ttg_chol_dpotri_tiled(nullptr, N, NB, tiles, info));
        if (sync()) return EXIT_FAILURE;
        const auto e = clk::now();
        if (t > 0) {
          line("POTRF+POTRI", "POTRF ", ms(b0, b1), flops_dpotrf(N));
          line("POTRF+POTRI", "POTRI ", ms(b1, e), flops_dpotri(N));
          line("POTRF+POTRI", "", ms(b0, e), flops_dpotrf(N) + flops_dpotri(N));
        }
      } else {
        CHECK_RC(ttg_chol_dtrtri_tiled(nullptr, N, NB, tiles, info));
        if (sync()) return EXIT_FAILURE;
        const auto b2 = clk::now();
        CHECK_RC(ttg_chol_dlauum_tiled(nullptr, N, NB, tiles));
        if (sync()) return EXIT_FAILURE;
        const auto e = clk::now();
        if (t > 0) {
          line("POTRF+TRTRI+LAUUM", "POTRF ", ms(b0, b1), flops_dpotrf(N));
          line("POTRF+TRTRI+LAUUM", "TRTRI ", ms(b1, b2), flops_dtrtri(N));
          line("POTRF+TRTRI+LAUUM", "LAUUM ", ms(b2, e), flops_dpotri(N) - flops_dtrtri(N));
          line("POTRF+TRTRI+LAUUM", "", ms(b0, e), flops_dpotrf(N) + flops_dpotri(N));
        }
      }
    }
    int h_info = 0;
    if (hipMemcpy(&h_info, info, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return EXIT_FAILURE;
    if (h_info != 0) {
      std::cout << "POINV failed: info = " << h_info << std::endl;
      ret = EXIT_FAILURE;
      break;
    }
  }
  if (ret == EXIT_SUCCESS && !A0.empty()) {
    std::vector<double> X(nel);
    if (hipMemcpy(X.data(), tiles, nel * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return EXIT_FAILURE;
    if (check_dpoinv(A0, X, N, NB) != 0) ret = EXIT_FAILURE;
  }
  (void)hipFree(tiles);
  (void)hipFree(info);
  return ret;
}
