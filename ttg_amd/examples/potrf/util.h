// Tile-norm consistency check of DEBUG_TILES_VALUES builds
// (examples/potrf/util.h:5-10): two 2-norms of one tile agree when they
// differ by at most 1e-12 of the larger.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <iostream>
#include <sstream>

#include "ttg.h"

inline bool check_norm(double expected, double actual) {
  return std::abs(expected - actual) <= std::max(std::abs(expected), std::abs(actual)) * 1e-12;
}

// 2-norm of n contiguous values on the host, summed in extended precision
// (the first norm of a tile that no task has produced yet).
template <typename T>
T host_nrm2(std::size_t n, const T *x) {
  long double s = 0.0L;
  for (std::size_t i = 0; i < n; ++i) s += (long double)x[i] * (long double)x[i];
  return (T)std::sqrt(s);
}

// Abort the run when a task received a tile other than the one its producer
// sent (a race or an ownership error in the runtime).
template <typename Key>
void expect_norm(const char *task, const Key &key, const char *tile, double expected, double actual) {
  if (check_norm(expected, actual)) return;
  std::ostringstream m;
  m.precision(17);
  m << task << "(" << key << "): " << tile << " has norm " << actual << ", its producer sent " << expected;
  std::cerr << m.str() << std::endl;
  ttg::abort();
}
