// C entry point running the examples/potrf task graph on the MI355X runtime
// over a caller-provided host matrix (lower tile-packed, NB x NB tiles), for
// tests and Python callers.  libttg_graph.so.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "ttg_graph.h"
#include "lauum.h"
#include "plgsy.h"
#include "pmw.h"
#include "potrf.h"
#include "potri.h"
#include "trtri_L.h"
#include "result.h"
#include "ttg.h"

namespace {

// load -> graph -> result writer over the host tiles, then fence; the
// statistics and the info of a stopped graph as ttg_graph_dpotrf's
template <typename MakeGraph>
int run_tile_graph(int N, int NB, double *tiles, long *stats, int *info, MakeGraph &&make_graph) {
  if (N <= 0 || NB <= 0 || NB > TTG_CHOL_MAX_TILE || !tiles) return -1;
  ttg::initialize(1, nullptr, -1);
  ttg::clear_cancel();
  auto &rt = ttg::detail::Runtime::get();
  const long t0 = rt.tasks_run, r0 = rt.rounds, k0 = rt.kernel_launches;
  rt.max_round = 0;
  {
    MatrixT<double> A(N, NB, tiles);
    ttg::Edge<Key2, MatrixTile<double>> to_graph("To graph"), to_result("To Result");
    auto load = make_load_tt(A, to_graph, true);
    auto graph = make_graph(A, to_graph, to_result);
    auto result_ttg = make_result_ttg(A, to_result, true);
    ttg::make_graph_executable(load.get());
    load->invoke();
    ttg::execute();
    ttg::fence();
  }
  if (stats) {
    stats[0] = rt.tasks_run - t0;
    stats[1] = rt.rounds - r0;
    stats[2] = rt.kernel_launches - k0;
    stats[3] = rt.max_round;
  }
  if (info) *info = ttg::cancelled() ? (int)ttg::cancel_info() : 0;
  ttg::clear_cancel();
  return 0;
}

// FP32 copy of the lower tiles (padding included) and the row sums of |A|
// of the symmetric matrix they hold, over a few host threads (by tile
// columns, each with its own row sums)
void to_float_tiles(int N, int NB, const double *A, float *Af, std::vector<double> &rowsum) {
  const int nt = (N + NB - 1) / NB;
  const size_t per = (size_t)NB * NB;
  const int nth = std::max(1, std::min<int>(8, (int)std::thread::hardware_concurrency()));
  std::vector<std::vector<double>> part(nth, std::vector<double>(N, 0.0));
  std::vector<std::thread> th;
  for (int w = 0; w < nth; ++w)
    th.emplace_back([&, w] {
      std::vector<double> &rs = part[w];
      for (int n = w; n < nt; n += nth)
        for (int m = n; m < nt; ++m) {
          const size_t t = ((size_t)n * nt - (size_t)n * (n - 1) / 2 + (m - n)) * per;
          for (size_t e = 0; e < per; ++e) Af[t + e] = (float)A[t + e];
          const int mb = std::min(NB, N - m * NB), nb = std::min(NB, N - n * NB);
          for (int j = 0; j < nb; ++j)
            for (int i = (m == n ? j : 0); i < mb; ++i) {
              const double v = std::fabs(A[t + i + (size_t)j * NB]);
              const int gi = m * NB + i, gj = n * NB + j;
              rs[gi] += v;
              if (gi != gj) rs[gj] += v;
            }
        }
    });
  for (auto &t : th) t.join();
  rowsum.assign(N, 0.0);
  for (auto &p : part)
    for (int i = 0; i < N; ++i) rowsum[i] += p[i];
}

// the padding of the last tile row of a factor made by the graph (which
// never touches it): zero, with a unit diagonal in the diagonal tile, as
// the tiled solves expect
void pad_factor(int N, int NB, float *L) {
  const int nt = (N + NB - 1) / NB, r = N - (nt - 1) * NB;
  if (r == NB) return;
  const size_t per = (size_t)NB * NB;
  for (int n = 0; n < nt; ++n) {
    float *t = L + ((size_t)n * nt - (size_t)n * (n - 1) / 2 + (nt - 1 - n)) * per;
    for (int j = 0; j < NB; ++j)
      for (int i = 0; i < NB; ++i) {
        const bool pad = i >= r || (n == nt - 1 && j >= r);
        if (pad) t[i + (size_t)j * NB] = (n == nt - 1 && i == j) ? 1.0f : 0.0f;
      }
  }
}

}  // namespace

extern "C" {

// Mixed-precision SPD solve whose factorization runs through the task graph:
// the potrf graph over FP32 tiles (device tasks on the S tile kernels), then
// iterative refinement in FP64 on the device (FP64 residual of the original
// matrix, FP32-factor solves) until max_j ||r_j||/(||A|| ||x_j|| + ||b_j||)
// <= tol or max_iter corrections.  A: host FP64 lower tiles (not modified);
// B, X: host N x nrhs column-major.  *iters: corrections applied; *berr: the
// final backward error; *info: 0 or LAPACK's info of the FP32 factorization
// (then nothing is solved).  Returns 0, -1 on a bad argument, -2 on a device
// error.
int ttg_graph_dposv_mixed(int N, int NB, const double *A, int nrhs, const double *B, int ldb, double *X, int ldx,
                          double tol, int max_iter, int *iters, double *berr, int *info) {
  if (N <= 0 || NB <= 0 || NB > TTG_CHOL_MAX_TILE || NB % 32 != 0 || nrhs <= 0 || ldb < N || ldx < N || !A || !B ||
      !X || max_iter < 0)
    return -1;
  ttg::initialize(1, nullptr, -1);
  ttg::clear_cancel();
  const int nt = (N + NB - 1) / NB;
  const size_t nel = (size_t)nt * (nt + 1) / 2 * NB * NB;
  std::vector<double> rowsum;
  MatrixT<float> Af(N, NB);
  to_float_tiles(N, NB, A, Af.tile_ptr(0, 0), rowsum);
  const double anorm = *std::max_element(rowsum.begin(), rowsum.end());
  {
    ttg::Edge<Key2, MatrixTile<float>> to_potrf("To POTRF"), to_result("To Result");
    auto load = make_load_tt(Af, to_potrf, true);
    auto potrf_ttg = potrf::make_potrf_ttg(Af, to_potrf, to_result, true);
    auto result_ttg = make_result_ttg(Af, to_result, true);
    ttg::make_graph_executable(load.get());
    load->invoke();
    ttg::execute();
    ttg::fence();
  }
  const int finfo = ttg::cancelled() ? (int)ttg::cancel_info() : 0;
  ttg::clear_cancel();
  if (info) *info = finfo;
  if (iters) *iters = 0;
  if (berr) *berr = NAN;
  if (finfo) return 0;
  pad_factor(N, NB, Af.tile_ptr(0, 0));

  // refinement on the device
  double *dA = nullptr, *dX = nullptr, *dB = nullptr, *dR = nullptr;
  float *dL = nullptr;
  hipStream_t s = nullptr;
  const size_t vec = (size_t)N * nrhs;
  int rc = 0;
  auto ok = [&](hipError_t e) {
    if (e != hipSuccess && !rc) rc = -2;
    return e == hipSuccess;
  };
  auto okc = [&](int e) {
    if (e != TTG_CHOL_OK && !rc) rc = -2;
    return e == TTG_CHOL_OK;
  };
  std::vector<double> Xh(vec), Bh(vec), Rh(vec);
  for (int j = 0; j < nrhs; ++j) std::copy(B + (size_t)j * ldb, B + (size_t)j * ldb + N, Bh.begin() + (size_t)j * N);
  if (ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) && ok(hipMalloc(&dA, nel * sizeof(double))) &&
      ok(hipMalloc(&dL, nel * sizeof(float))) && ok(hipMalloc(&dX, vec * sizeof(double))) &&
      ok(hipMalloc(&dB, vec * sizeof(double))) && ok(hipMalloc(&dR, vec * sizeof(double))) &&
      ok(hipMemcpyAsync(dA, A, nel * sizeof(double), hipMemcpyHostToDevice, s)) &&
      ok(hipMemcpyAsync(dL, Af.tile_ptr(0, 0), nel * sizeof(float), hipMemcpyHostToDevice, s)) &&
      ok(hipMemcpyAsync(dB, Bh.data(), vec * sizeof(double), hipMemcpyHostToDevice, s)) &&
      ok(hipMemcpyAsync(dX, Bh.data(), vec * sizeof(double), hipMemcpyHostToDevice, s)) &&
      okc(ttg_chol_spotrs_tiled(s, N, NB, dL, nrhs, dX, N)) &&
      ok(hipMemcpyAsync(Xh.data(), dX, vec * sizeof(double), hipMemcpyDeviceToHost, s)) && ok(hipStreamSynchronize(s))) {
    int it = 0;
    for (;;) {
      if (!okc(ttg_chol_dresidual_tiled(s, N, NB, dA, nrhs, dX, N, dB, N, dR, N)) ||
          !ok(hipMemcpyAsync(Rh.data(), dR, vec * sizeof(double), hipMemcpyDeviceToHost, s)) ||
          !ok(hipStreamSynchronize(s)))
        break;
      double e = 0.0;
      for (int j = 0; j < nrhs; ++j) {
        double rn = 0.0, xn = 0.0, bn = 0.0;
        for (int i = 0; i < N; ++i) {
          const size_t k = (size_t)j * N + i;
          rn = std::max(rn, std::fabs(Rh[k]));
          xn = std::max(xn, std::fabs(Xh[k]));
          bn = std::max(bn, std::fabs(Bh[k]));
          if (std::isnan(Rh[k])) rn = NAN;
        }
        const double ej = rn / (anorm * xn + bn);
        e = std::isnan(ej) || std::isnan(e) ? NAN : std::max(e, ej);
      }
      if (berr) *berr = e;
      if (!(e > tol) || it >= max_iter) break;  // converged (or NaN: stop), or out of steps
      // X += A32^{-1} R
      if (!okc(ttg_chol_spotrs_tiled(s, N, NB, dL, nrhs, dR, N)) ||
          !ok(hipMemcpyAsync(Rh.data(), dR, vec * sizeof(double), hipMemcpyDeviceToHost, s)) ||
          !ok(hipStreamSynchronize(s)))
        break;
      for (size_t k = 0; k < vec; ++k) Xh[k] += Rh[k];
      if (!ok(hipMemcpyAsync(dX, Xh.data(), vec * sizeof(double), hipMemcpyHostToDevice, s))) break;
      ++it;
    }
    if (iters) *iters = it;
  }
  for (int j = 0; j < nrhs; ++j) std::copy(Xh.begin() + (size_t)j * N, Xh.begin() + (size_t)(j + 1) * N, X + (size_t)j * ldx);
  if (s) (void)hipStreamSynchronize(s);
  (void)hipFree(dA);
  (void)hipFree(dL);
  (void)hipFree(dX);
  (void)hipFree(dB);
  (void)hipFree(dR);
  if (s) (void)hipStreamDestroy(s);
  (void)ttg_chol_release_plans();  // the solve plans hold pointers into the freed tiles
  return rc;
}


// Factor the N x N SPD matrix held in `tiles` (host, lower tile-packed) in
// place with the TTG graph; stats (optional, 4 longs): tasks run, scheduling
// rounds, kernel launches, largest round.  *info (optional) receives 0, or
// LAPACK's info of a tile that is not positive definite: the graph stops
// there and the tiles are left partly factored.  Returns 0, or -1 on a bad
// argument.
int ttg_graph_dpotrf(int N, int NB, double *tiles, long *stats, int *info) {
  if (N <= 0 || NB <= 0 || NB > TTG_CHOL_MAX_TILE || !tiles) return -1;
  ttg::initialize(1, nullptr, -1);  // host workers: TTG_NUM_THREADS, else the runtime's default
  ttg::clear_cancel();
  auto &rt = ttg::detail::Runtime::get();
  const long t0 = rt.tasks_run, r0 = rt.rounds, k0 = rt.kernel_launches;
  rt.max_round = 0;
  {
    MatrixT<double> A(N, NB, tiles);
    ttg::Edge<Key2, MatrixTile<double>> topotrf("To POTRF"), toresult("To Result");
    auto load = make_load_tt(A, topotrf, true);
    auto potrf_ttg = potrf::make_potrf_ttg(A, topotrf, toresult, true);
    auto result_ttg = make_result_ttg(A, toresult, true);
    ttg::make_graph_executable(load.get());
    load->invoke();
    ttg::execute();
    ttg::fence();
  }
  if (stats) {
    stats[0] = rt.tasks_run - t0;
    stats[1] = rt.rounds - r0;
    stats[2] = rt.kernel_launches - k0;
    stats[3] = rt.max_round;
  }
  if (info) *info = ttg::cancelled() ? (int)ttg::cancel_info() : 0;
  ttg::clear_cancel();
  return 0;
}


// L := L^{-1} on the lower tiles with the TRTRI graph (trtri_L.h).  *info: 0
// or LAPACK's info of an exactly singular diagonal.
int ttg_graph_dtrtri(int N, int NB, double *tiles, long *stats, int *info) {
  return run_tile_graph(N, NB, tiles, stats, info, [](auto &A, auto &in, auto &out) {
    return trtri_lower::make_trtri_ttg(A, in, out, true);
  });
}

// A := L^T L on the lower tiles with the LAUUM graph (lauum.h).
int ttg_graph_dlauum(int N, int NB, double *tiles, long *stats) {
  return run_tile_graph(N, NB, tiles, stats, nullptr, [](auto &A, auto &in, auto &out) {
    return lauum::make_lauum_ttg(A, in, out, true);
  });
}

// A^{-1} from the factor L (lower tiles) with the POTRI graph (potri.h).
int ttg_graph_dpotri(int N, int NB, double *tiles, long *stats, int *info) {
  return run_tile_graph(N, NB, tiles, stats, info, [](auto &A, auto &in, auto &out) {
    return potri::make_potri_ttg(A, in, out, true);
  });
}

// Generate PLGSY(N, bump, seed) into `tiles` (host) with the device-task
// PLGSY graph.
int ttg_graph_dplgsy(int N, int NB, double bump, unsigned long seed, double *tiles) {
  if (N <= 0 || NB <= 0 || !tiles) return -1;
  ttg::initialize(1, nullptr, -1);  // host workers: TTG_NUM_THREADS, else the runtime's default
  MatrixT<double> A(N, NB, tiles);
  ttg::Edge<Key2, void> startup("startup");
  ttg::Edge<Key2, MatrixTile<double>> result("To result");
  auto init_tt = ttg::make_tt<void>(
      [&](std::tuple<ttg::Out<Key2, void>> &out) {
        for (int i = 0; i < (int)A.rows(); i++)
          for (int j = 0; j <= i; j++) ttg::sendk<0>(Key2{i, j}, out);
      },
      ttg::edges(), ttg::edges(startup), "Startup Trigger");
  auto plgsy_ttg = make_plgsy_ttg(A, (unsigned long)bump, seed, startup, result, true);
  auto result_ttg = make_result_ttg(A, result, true);
  init_tt->invoke();
  ttg::fence();
  return 0;
}

}  // extern "C"
