// Build recipe helper for oracle/_ref: exposes the reference's own generator
// headers (examples/potrf/core_plgsy.h, random.h, util.h, compiled from where
// they lie in the reference tree) through a C ABI so the tests can compare the
// oracle's restatement against them.  Test infrastructure only.
#include <algorithm>
#include <cmath>
#include <complex>
#include <type_traits>

#include "core_plgsy.h"
#include "util.h"

extern "C" {
void ref_core_plgsy(double bump, int m, int n, double *A, int lda, int gM, int m0, int n0,
                    unsigned long long seed) {
  CORE_plgsy<double>(bump, m, n, A, lda, gM, m0, n0, seed);
}
unsigned long long ref_rnd64_jump(unsigned long long n, unsigned long long seed) {
  return Rnd64_jump(n, seed);
}
int ref_check_norm(double expected, double actual) { return check_norm(expected, actual) ? 1 : 0; }
}
