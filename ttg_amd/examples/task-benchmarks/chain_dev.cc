// Device-task overhead of the runtime: a chain of dependent device tasks
// (the task benchmark examples/task-benchmarks/chain-ttg-dev.cc).  Task k
// selects the buffers of F values (F = 1, 2, 4 or 8 flows) and forwards them
// to task k + 1; with F = 0 the task ids flow through keys alone and the
// task selects nothing.  The chain is run once untimed and once timed, and
// the reference's lines are printed:
//   # of tasks = <n>
//   time elapsed (microseconds) = <t>, avg <t / n>
// and, between them, "# of tasks run = <count>" from the runtime's own
// count of the timed run (n + 1 when the chain runs to its end).
// Usage: chain_ttg_dev [F] [-n <tasks>]   (defaults: F = 0, 100000 tasks)
// No task launches a kernel, so the time is the runtime's own: a scheduling
// round per task (its launch group and completion event), the buffer
// residency bookkeeping and the deferred sends.
#include <array>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <tuple>
#include <utility>
#include <vector>

#include "ttg.h"

namespace {

int g_ntasks = 100000;

// a value carrying one device buffer of a single int
struct Token {
  ttg::Buffer<int> b;
  Token() : b(1) {}
  Token(Token &&) = default;
};
std::vector<ttg::detail::BufferBase *> ttg_buffers(Token &t) { return {&t.b}; }

template <std::size_t>
using TokenOf = Token;

// task body for F = sizeof...(I) value flows
template <typename Outs, std::size_t... I>
struct ValueChain {
  ttg::device::Task operator()(const int &key, TokenOf<I> &&...v, Outs &out) const {
    co_await ttg::device::select(v.b...);
    if (key < g_ntasks) co_await ttg::device::forward(ttg::device::send<I>(key + 1, std::move(v), out)...);
  }
};

// task body for F = 0: ids in the keys only
struct KeyChain {
  ttg::device::Task operator()(const int &key, std::tuple<ttg::Out<int, void>> &out) const {
    co_await ttg::device::select();
    if (key < g_ntasks) co_await ttg::device::forward(ttg::device::sendk<0>(key + 1, out));
  }
};

long g_tasks_run = 0;  // tasks the runtime ran in the timed run

template <typename Init>
double run_twice(Init &init) {
  init->invoke();
  ttg::execute();
  ttg::fence();
  const long before = ttg::detail::Runtime::get().tasks_run;
  const auto t0 = std::chrono::high_resolution_clock::now();
  init->invoke();
  ttg::execute();
  ttg::fence();
  const auto t1 = std::chrono::high_resolution_clock::now();
  g_tasks_run = ttg::detail::Runtime::get().tasks_run - before;
  return std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count();
}

template <std::size_t... I>
double bench_values(std::index_sequence<I...>) {
  using Outs = std::tuple<ttg::Out<int, TokenOf<I>>...>;
  std::array<ttg::Edge<int, Token>, sizeof...(I)> to_first, to_next;
  auto init = ttg::make_tt([](Outs &out) { (ttg::send<I>(0, Token{}, out), ...); }, ttg::edges(),
                           ttg::edges(to_first[I]...), "init");
  auto next = ttg::make_tt<ttg::ExecutionSpace::HIP>(ValueChain<Outs, I...>{},
                                                     ttg::edges(ttg::fuse(to_first[I], to_next[I])...),
                                                     ttg::edges(to_next[I]...), "next");
  std::cout << "Graph " << sizeof...(I) << " is connected.\n";
  return run_twice(init);
}

double bench_keys() {
  ttg::Edge<int, void> to_first, to_next;
  auto init = ttg::make_tt([](std::tuple<ttg::Out<int, void>> &out) { ttg::sendk<0>(0, out); }, ttg::edges(),
                           ttg::edges(to_first), "init");
  auto next = ttg::make_tt<ttg::ExecutionSpace::HIP>(KeyChain{}, ttg::edges(ttg::fuse(to_first, to_next)),
                                                     ttg::edges(to_next), "next");
  std::cout << "Graph 0 is connected.\n";
  return run_twice(init);
}

}  // namespace

int main(int argc, char **argv) {
  int flows = 0;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "-n") && i + 1 < argc)
      g_ntasks = std::atoi(argv[++i]);
    else
      flows = std::atoi(argv[i]);
  }
  ttg::initialize(argc, argv, -1);
  double us = -1.0;
  switch (flows) {
    case 0: us = bench_keys(); break;
    case 1: us = bench_values(std::make_index_sequence<1>{}); break;
    case 2: us = bench_values(std::make_index_sequence<2>{}); break;
    case 4: us = bench_values(std::make_index_sequence<4>{}); break;
    case 8: us = bench_values(std::make_index_sequence<8>{}); break;
    default: std::cout << "Unsupported number of flows: " << flows << std::endl;
  }
  if (us >= 0.0) {
    // the key-only chain runs tasks 0..n, the value chains likewise
    std::cout << "# of tasks = " << g_ntasks << std::endl;
    // the chain's own count of the tasks that ran (keys 0 .. n; the init
    // body runs at invoke(), outside the runtime)
    std::cout << "# of tasks run = " << g_tasks_run << std::endl;
    std::cout << "time elapsed (microseconds) = " << us << ", avg " << us / (double)g_ntasks << std::endl;
  }
  ttg::finalize();
  return us >= 0.0 ? EXIT_SUCCESS : EXIT_FAILURE;
}
