// Small task graphs that pin the runtime's data semantics (ttg.h), each
// printing one line "<name>: PASS" or "<name>: FAIL <why>"; the exit status
// is non-zero when any fails.
//   reader_writer_defer  one broadcast tile to a device reader (const &) and
//   reader_writer_copy   a device writer (&&, factors it in place), the writer
//                        first by priority: the reader must see the tile as
//                        sent.  With set_defer_writer(true) the writer waits
//                        for the reader (a later round, same buffer); without
//                        it the writer works on a copy of its own
//                        (register_data_copy, parsec/ttg.h:958-1060).
//   host_route           device writer -> host router (const &, never reads
//                        the data) -> device task -> host writer (&&, edits the
//                        host copy) -> device reader -> host sink: the router
//                        costs no push-out, the host writer's edit reaches
//                        the device reader, kernels launched after a wait see
//                        the task's data, and the sink reads the edited tile.
//   cancel_chain         a chain of device tasks whose task 37 stops the
//   cancel_fanout        graph with ttg::cancel after a wait; a fan-out whose
//                        first task stops it: nothing after the stop runs,
//                        fence() returns, the info is kept, and the runtime
//                        runs the next graph in full.
//   fanout_host          64 device tasks factor shifted tiles, 64 host tasks
//                        (on the worker threads when -c > 1) check and edit
//                        them, 64 device tasks read the edits: every push-out
//                        and every send recorded on a worker arrives.
// Usage: runtime_checks [-c threads]   (default 1: host tasks inline)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "../potrf/matrixtile.h"
#include "ttg.h"
#include "ttg_chol.h"

namespace {

using Tile = MatrixTile<double>;
constexpr int kN = 64;
int g_failures = 0;

void report(const std::string &name, const std::string &why) {
  if (why.empty()) {
    std::cout << name << ": PASS" << std::endl;
  } else {
    std::cout << name << ": FAIL " << why << std::endl;
    ++g_failures;
  }
}

void check_launch(int rc) {
  if (rc != TTG_CHOL_OK) ttg::detail::fatal(std::string("tile kernel: ") + ttg_chol_last_error());
}

// an SPD tile: 1 + i + j off the diagonal scaled down, 2 n on it
std::vector<double> spd_values(int n) {
  std::vector<double> a((std::size_t)n * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) a[i + (std::size_t)j * n] = i == j ? 2.0 * n : 1.0 / (1.0 + i + j);
  return a;
}

Tile make_tile(const std::vector<double> &a, int n) {
  Tile t(n, n, n);
  std::copy(a.begin(), a.end(), t.data());
  return t;
}

double host_norm(const double *x, std::size_t n) {
  long double s = 0.0L;
  for (std::size_t i = 0; i < n; ++i) s += (long double)x[i] * x[i];
  return (double)std::sqrt(s);
}

bool close(double a, double b, double rel) { return std::abs(a - b) <= rel * std::max(std::abs(a), std::abs(b)); }

// max |L L^T - A| over the lower triangle, relative to max |A|
double factor_error(const std::vector<double> &L, const std::vector<double> &A, int n) {
  double err = 0.0, amax = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = j; i < n; ++i) {
      double s = 0.0;
      for (int k = 0; k <= j; ++k) s += L[i + (std::size_t)k * n] * L[j + (std::size_t)k * n];
      err = std::max(err, std::abs(s - A[i + (std::size_t)j * n]));
      amax = std::max(amax, std::abs(A[i + (std::size_t)j * n]));
    }
  return err / amax;
}

void reader_writer(bool defer) {
  auto &rt = ttg::detail::Runtime::get();
  const std::vector<double> a = spd_values(kN);
  const double a_norm = host_norm(a.data(), a.size());
  double reader_norm = -1.0;
  long reader_round = -1, writer_round = -1;
  const void *sent_host = nullptr, *writer_host = nullptr;
  int writer_info = -1;
  std::vector<double> factor;

  ttg::Edge<int, Tile> to_reader("to reader"), to_writer("to writer"), to_sink("to sink");
  auto src = ttg::make_tt(
      [&](std::tuple<ttg::Out<int, Tile>, ttg::Out<int, Tile>> &out) {
        Tile t = make_tile(a, kN);
        sent_host = t.buffer().host;
        ttg::broadcast<0, 1>(std::make_tuple(0, 0), std::move(t), out);
      },
      ttg::edges(), ttg::edges(to_reader, to_writer), "source");
  auto reader = ttg::make_tt<ttg::ExecutionSpace::HIP>(
      [&](const int &, const Tile &t, std::tuple<> &) -> ttg::device::Task {
        double nrm = -1.0;
        auto s = ttg::make_scratch(&nrm, ttg::scope::Allocate);
        co_await ttg::device::select(t.buffer(), s);
        check_launch(ttg_chol_dnrm2(ttg::device::current_stream(), (int)t.size(), t.buffer().current_device_ptr(),
                                    s.device_ptr()));
        reader_round = rt.rounds;
        co_await ttg::device::wait(s);
        reader_norm = nrm;
      },
      ttg::edges(to_reader), ttg::edges(), "reader");
  auto writer = ttg::make_tt<ttg::ExecutionSpace::HIP>(
      [&](const int &k, Tile &&t, std::tuple<ttg::Out<int, Tile>> &out) -> ttg::device::Task {
        int info = -1;
        auto s = ttg::make_scratch(&info, ttg::scope::Allocate);
        co_await ttg::device::select(t.buffer(), s);
        check_launch(ttg_chol_dpotrf_tile(ttg::device::current_stream(), (int)t.cols(),
                                          t.buffer().current_device_ptr(), (int)t.lda(), s.device_ptr(), 0));
        writer_round = rt.rounds;
        writer_host = t.buffer().host;
        co_await ttg::device::wait(s);
        writer_info = info;
        co_await ttg::device::send<0>(k, std::move(t), out);
      },
      ttg::edges(to_writer), ttg::edges(to_sink), "writer");
  auto sink = ttg::make_tt([&](const int &, Tile &&t, std::tuple<> &) { factor.assign(t.data(), t.data() + t.size()); },
                           ttg::edges(to_sink), ttg::edges(), "sink");
  writer->set_defer_writer(defer);
  // the writer is taken from the ready queue first
  writer->set_priomap([](const int &) { return 10; });
  reader->set_priomap([](const int &) { return 1; });
  src->invoke();
  ttg::fence();

  std::ostringstream why;
  if (!close(reader_norm, a_norm, 1e-12))
    why << "the reader saw norm " << reader_norm << ", the tile sent has norm " << a_norm << "; ";
  if (writer_info != 0) why << "writer info " << writer_info << "; ";
  if (factor.size() != a.size() || factor_error(factor, a, kN) > 1e-13) why << "the writer's factor is wrong; ";
  if (defer) {
    if (writer_round <= reader_round) why << "the deferred writer ran in round " << writer_round << " <= the reader's; ";
    if (writer_host != sent_host) why << "the deferred writer got a copy; ";
  } else if (writer_host == sent_host) {
    why << "the writer mutated the shared tile instead of a copy; ";
  }
  report(defer ? "reader_writer_defer" : "reader_writer_copy", why.str());
}

void host_route() {
  auto &rt = ttg::detail::Runtime::get();
  const std::vector<double> a = spd_values(kN);
  double norm_after_wait = -1.0, norm_q = -1.0, norm_r = -1.0, norm_L = -1.0;
  long push_outs_router = -1;
  std::vector<double> sunk;
  long p0 = 0;

  ttg::Edge<int, Tile> e_p("to P"), e_route("to router"), e_q("to Q"), e_w("to host writer"), e_r("to R"),
      e_sink("to sink");
  auto src = ttg::make_tt([&](std::tuple<ttg::Out<int, Tile>> &out) { ttg::send<0>(0, make_tile(a, kN), out); },
                          ttg::edges(), ttg::edges(e_p), "source");
  // P: factor, wait for info, then a norm launched after the wait
  auto P = ttg::make_tt<ttg::ExecutionSpace::HIP>(
      [&](const int &k, Tile &&t, std::tuple<ttg::Out<int, Tile>> &out) -> ttg::device::Task {
        int info = -1;
        double nrm = -1.0;
        auto si = ttg::make_scratch(&info, ttg::scope::Allocate);
        auto sn = ttg::make_scratch(&nrm, ttg::scope::Allocate);
        co_await ttg::device::select(t.buffer(), si, sn);
        check_launch(ttg_chol_dpotrf_tile(ttg::device::current_stream(), (int)t.cols(),
                                          t.buffer().current_device_ptr(), (int)t.lda(), si.device_ptr(), 0));
        co_await ttg::device::wait(si);
        if (info != 0) ttg::detail::fatal("host_route: the test tile is not SPD");
        check_launch(ttg_chol_dnrm2(ttg::device::current_stream(), (int)t.size(), t.buffer().current_device_ptr(),
                                    sn.device_ptr()));
        co_await ttg::device::wait(sn);
        norm_after_wait = nrm;
        co_await ttg::device::send<0>(k, std::move(t), out);
      },
      ttg::edges(e_p), ttg::edges(e_route), "P");
  auto router = ttg::make_tt(
      [&](const int &k, const Tile &t, std::tuple<ttg::Out<int, Tile>> &out) {
        p0 = rt.push_outs;
        ttg::send<0>(k, t, out);
        push_outs_router = rt.push_outs - p0;
      },
      ttg::edges(e_route), ttg::edges(e_q), "router");
  auto Q = ttg::make_tt<ttg::ExecutionSpace::HIP>(
      [&](const int &k, Tile &&t, std::tuple<ttg::Out<int, Tile>> &out) -> ttg::device::Task {
        double nrm = -1.0;
        auto sn = ttg::make_scratch(&nrm, ttg::scope::Allocate);
        co_await ttg::device::select(t.buffer(), sn);
        check_launch(ttg_chol_dnrm2(ttg::device::current_stream(), (int)t.size(), t.buffer().current_device_ptr(),
                                    sn.device_ptr()));
        co_await ttg::device::wait(sn);
        norm_q = nrm;
        co_await ttg::device::send<0>(k, std::move(t), out);
      },
      ttg::edges(e_q), ttg::edges(e_w), "Q");
  auto hw = ttg::make_tt(
      [&](const int &k, Tile &&t, std::tuple<ttg::Out<int, Tile>> &out) {
        norm_L = host_norm(t.data(), t.size());
        t.data()[0] += 1.0;
        ttg::send<0>(k, std::move(t), out);
      },
      ttg::edges(e_w), ttg::edges(e_r), "host writer");
  auto R = ttg::make_tt<ttg::ExecutionSpace::HIP>(
      [&](const int &k, const Tile &t, std::tuple<ttg::Out<int, Tile>> &out) -> ttg::device::Task {
        double nrm = -1.0;
        auto sn = ttg::make_scratch(&nrm, ttg::scope::Allocate);
        co_await ttg::device::select(t.buffer(), sn);
        check_launch(ttg_chol_dnrm2(ttg::device::current_stream(), (int)t.size(), t.buffer().current_device_ptr(),
                                    sn.device_ptr()));
        co_await ttg::device::wait(sn);
        norm_r = nrm;
        co_await ttg::device::send<0>(k, t, out);
      },
      ttg::edges(e_r), ttg::edges(e_sink), "R");
  auto sink = ttg::make_tt([&](const int &, const Tile &t, std::tuple<> &) { sunk.assign(t.data(), t.data() + t.size()); },
                           ttg::edges(e_sink), ttg::edges(), "sink");
  const long before = rt.push_outs;
  src->invoke();
  ttg::fence();
  const long total = rt.push_outs - before;

  std::ostringstream why;
  if (push_outs_router != 0) why << "the router caused " << push_outs_router << " push-outs; ";
  // only the host writer's first data() copies the tile back
  if (total != 1) why << total << " push-outs in all, expected 1; ";
  if (!(norm_after_wait > 0) || norm_after_wait != norm_q)
    why << "norm after the wait " << norm_after_wait << " differs from the next task's " << norm_q << "; ";
  if (!close(norm_q, norm_L, 1e-12)) why << "device norm of L " << norm_q << " vs host " << norm_L << "; ";
  const double want_r = sunk.empty() ? -1.0 : host_norm(sunk.data(), sunk.size());
  if (!close(norm_r, want_r, 1e-12)) why << "the device reader saw norm " << norm_r << ", the host edit gives " << want_r << "; ";
  if (close(norm_r, norm_q, 1e-15)) why << "the host writer's edit did not reach the device; ";
  report("host_route", why.str());
}

void cancel_chain() {
  auto &rt = ttg::detail::Runtime::get();
  constexpr int n = 100;
  int stop_at = 37, ran = 0;
  ttg::Edge<int, void> to_first("first"), to_next("next");
  auto init = ttg::make_tt([](std::tuple<ttg::Out<int, void>> &out) { ttg::sendk<0>(0, out); }, ttg::edges(),
                           ttg::edges(to_first), "init");
  auto step = ttg::make_tt<ttg::ExecutionSpace::HIP>(
      [&](const int &k, std::tuple<ttg::Out<int, void>> &out) -> ttg::device::Task {
        int flag = 0;
        auto s = ttg::make_scratch(&flag, ttg::scope::SyncIn);
        co_await ttg::device::select(s);
        co_await ttg::device::wait(s);
        ++ran;
        if (k == stop_at) {
          ttg::cancel(1000 + k);
          co_return;
        }
        if (k + 1 < n) co_await ttg::device::forward(ttg::device::sendk<0>(k + 1, out));
      },
      ttg::edges(ttg::fuse(to_first, to_next)), ttg::edges(to_next), "step");
  std::ostringstream why;
  ttg::clear_cancel();
  const long t0 = rt.tasks_run;
  init->invoke();
  ttg::fence();
  if (!ttg::cancelled()) why << "not cancelled; ";
  if (ttg::cancel_info() != 1037) why << "info " << ttg::cancel_info() << "; ";
  if (ran != stop_at + 1) why << ran << " tasks ran, expected " << stop_at + 1 << "; ";
  ttg::clear_cancel();
  // the next graph runs in full
  stop_at = -1;
  ran = 0;
  init->invoke();
  ttg::fence();
  if (ttg::cancelled()) why << "second run cancelled; ";
  if (ran != n) why << "second run: " << ran << " tasks ran, expected " << n << "; ";
  (void)t0;
  report("cancel_chain", why.str());
}

void cancel_fanout() {
  constexpr int n = 50;
  int fan_ran = 0, next_ran = 0;
  ttg::Edge<int, void> to_fan("fan"), to_next("next");
  auto init = ttg::make_tt(
      [](std::tuple<ttg::Out<int, void>> &out) {
        for (int k = 0; k < n; ++k) ttg::sendk<0>(k, out);
      },
      ttg::edges(), ttg::edges(to_fan), "init");
  auto fan = ttg::make_tt<ttg::ExecutionSpace::HIP>(
      [&](const int &k, std::tuple<ttg::Out<int, void>> &out) -> ttg::device::Task {
        co_await ttg::device::select();
        ++fan_ran;
        if (k == 0) ttg::cancel(7);
        co_await ttg::device::forward(ttg::device::sendk<0>(k, out));
      },
      ttg::edges(to_fan), ttg::edges(to_next), "fan");
  auto next = ttg::make_tt<ttg::ExecutionSpace::HIP>(
      [&](const int &, std::tuple<> &) -> ttg::device::Task {
        co_await ttg::device::select();
        ++next_ran;
      },
      ttg::edges(to_next), ttg::edges(), "next");
  fan->set_priomap([](const int &k) { return n - k; });
  ttg::clear_cancel();
  init->invoke();
  ttg::fence();
  std::ostringstream why;
  if (!ttg::cancelled() || ttg::cancel_info() != 7) why << "cancel state " << ttg::cancelled() << "/" << ttg::cancel_info() << "; ";
  // every fan task started in the stopping round finishes; none after it starts
  if (fan_ran < 1 || fan_ran > n) why << fan_ran << " fan tasks ran; ";
  if (next_ran != 0) why << next_ran << " tasks ran after the stop; ";
  ttg::clear_cancel();
  report("cancel_fanout", why.str());
}

void fanout_host() {
  constexpr int n = 64;
  const std::vector<double> a = spd_values(kN);
  std::vector<double> err(n, -1.0), edited_norm(n, -1.0), device_norm(n, -1.0);
  std::vector<int> infos(n, -1);
  ttg::Edge<int, Tile> e_p("to P"), e_h("to host"), e_r("to R");
  auto src = ttg::make_tt(
      [&](std::tuple<ttg::Out<int, Tile>> &out) {
        for (int k = 0; k < n; ++k) {
          std::vector<double> ak = a;
          for (int i = 0; i < kN; ++i) ak[i + (std::size_t)i * kN] += k;
          ttg::send<0>(k, make_tile(ak, kN), out);
        }
      },
      ttg::edges(), ttg::edges(e_p), "source");
  auto P = ttg::make_tt<ttg::ExecutionSpace::HIP>(
      [&](const int &k, Tile &&t, std::tuple<ttg::Out<int, Tile>> &out) -> ttg::device::Task {
        int info = -1;
        auto si = ttg::make_scratch(&info, ttg::scope::Allocate);
        co_await ttg::device::select(t.buffer(), si);
        check_launch(ttg_chol_dpotrf_tile(ttg::device::current_stream(), (int)t.cols(),
                                          t.buffer().current_device_ptr(), (int)t.lda(), si.device_ptr(), 0));
        co_await ttg::device::wait(si);
        infos[k] = info;
        co_await ttg::device::send<0>(k, std::move(t), out);
      },
      ttg::edges(e_p), ttg::edges(e_h), "P");
  auto H = ttg::make_tt(
      [&](const int &k, Tile &&t, std::tuple<ttg::Out<int, Tile>> &out) {
        std::vector<double> L(t.data(), t.data() + t.size()), ak = a;
        for (int i = 0; i < kN; ++i) ak[i + (std::size_t)i * kN] += k;
        err[k] = factor_error(L, ak, kN);
        t.data()[1] = 100.0 + k;  // (1, 0): part of L
        edited_norm[k] = host_norm(t.data(), t.size());
        ttg::send<0>(k, std::move(t), out);
      },
      ttg::edges(e_h), ttg::edges(e_r), "host check");
  auto R = ttg::make_tt<ttg::ExecutionSpace::HIP>(
      [&](const int &k, const Tile &t, std::tuple<> &) -> ttg::device::Task {
        double nrm = -1.0;
        auto sn = ttg::make_scratch(&nrm, ttg::scope::Allocate);
        co_await ttg::device::select(t.buffer(), sn);
        check_launch(ttg_chol_dnrm2(ttg::device::current_stream(), (int)t.size(), t.buffer().current_device_ptr(),
                                    sn.device_ptr()));
        co_await ttg::device::wait(sn);
        device_norm[k] = nrm;
      },
      ttg::edges(e_r), ttg::edges(), "R");
  src->invoke();
  ttg::fence();
  std::ostringstream why;
  for (int k = 0; k < n; ++k) {
    if (infos[k] != 0) why << "P(" << k << ") info " << infos[k] << "; ";
    if (!(err[k] >= 0.0 && err[k] <= 1e-13)) why << "H(" << k << ") factor error " << err[k] << "; ";
    if (!close(device_norm[k], edited_norm[k], 1e-12))
      why << "R(" << k << ") saw norm " << device_norm[k] << ", the host edit gives " << edited_norm[k] << "; ";
  }
  report("fanout_host", why.str());
}

}  // namespace

int main(int argc, char **argv) {
  int threads = 1;
  for (int i = 1; i + 1 < argc; ++i)
    if (!std::strcmp(argv[i], "-c")) threads = std::atoi(argv[i + 1]);
  ttg::initialize(argc, argv, threads);
  std::cout << "threads " << ttg::detail::Runtime::get().num_threads() << std::endl;
  reader_writer(true);
  reader_writer(false);
  host_route();
  cancel_chain();
  cancel_fanout();
  fanout_host();
  ttg::finalize();
  std::cout << (g_failures ? "runtime checks FAILED" : "runtime checks passed") << std::endl;
  return g_failures ? EXIT_FAILURE : EXIT_SUCCESS;
}
