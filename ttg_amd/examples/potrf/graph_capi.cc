// C entry point running the examples/potrf task graph on the MI355X runtime
// over a caller-provided host matrix (lower tile-packed, NB x NB tiles), for
// tests and Python callers.  libttg_graph.so.
#include <cstring>

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

}  // namespace

extern "C" {

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
