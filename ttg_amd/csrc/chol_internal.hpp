// Internal helpers shared by the C-ABI translation units of libttg_chol.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/ttg_chol.h"

namespace ttgc {

// thread-local message of the last failed call
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();

inline int hip_fail(hipError_t e, const char *what) {
  set_error("%s: %s", what, hipGetErrorString(e));
  return TTG_CHOL_EHIP;
}

#define TTGC_HIP(...)                                           \
  do {                                                          \
    hipError_t _e = (__VA_ARGS__);                              \
    if (_e != hipSuccess) return ttgc::hip_fail(_e, #__VA_ARGS__); \
  } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Schedule knob `which` (TTG_CHOL_KNOB_*) for a plan being built: the value
// ttg_chol_set_knob gave, else the environment variable `env`, else `def`.
int plan_knob(int which, const char *env, int def);

// Two streams whose HSA queues split the CUs of the current device: `few` may
// only use CUs 0 .. want-1 of the device mask, `rest` the others (KFD deals
// the mask bits round-robin over the XCDs, so want = 8 is one CU per XCD).
// Work on `few` then never waits behind thousands of workgroups on `rest`.
// Returns false (and creates nothing) if the device is too small for the
// split or the runtime refuses the masks; the caller then uses plain streams.
inline bool make_split_streams(int want, hipStream_t *few, hipStream_t *rest) {
  int dev = 0, ncu = 0;
  if (want <= 0 || hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || want * 4 > ncu)
    return false;
  const int words = (ncu + 31) / 32;
  uint32_t mfew[16] = {0}, mrest[16] = {0};
  if (words > 16) return false;
  for (int c = 0; c < ncu; ++c) (c < want ? mfew : mrest)[c / 32] |= 1u << (c % 32);
  hipStream_t a = nullptr, b = nullptr;
  if (hipExtStreamCreateWithCUMask(&a, words, mfew) != hipSuccess ||
      hipExtStreamCreateWithCUMask(&b, words, mrest) != hipSuccess) {
    if (a) (void)hipStreamDestroy(a);
    if (b) (void)hipStreamDestroy(b);
    (void)hipGetLastError();
    return false;
  }
  *few = a;
  *rest = b;
  return true;
}

// Device buffer that grows on demand (never shrinks); used for descriptor
// staging and workspaces owned by the library.
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  hipError_t reserve(size_t n) {
    if (n <= bytes) return hipSuccess;
    if (ptr) {
      hipError_t e = hipFree(ptr);
      if (e != hipSuccess) return e;
      ptr = nullptr;
      bytes = 0;
    }
    size_t want = n < 4096 ? 4096 : n;
    hipError_t e = hipMalloc(&ptr, want);
    if (e == hipSuccess) bytes = want;
    return e;
  }
};

// Drop the cached plans and workspaces of one translation unit (device
// memory, streams, events, graphs).  The caller has synchronised the device
// and no other library call is in flight.
void release_tiled_plans();
void release_solve_plans();
void release_mixed_plans();
void release_poinv_plans();
void release_check_work();

// dst[i] := (double)src[i] for i < n (FP32 or FP64 tiles to FP64).
template <typename T>
hipError_t launch_widen(hipStream_t s, long n, const T *src, double *dst);

// *r := ||x||_2 of n contiguous doubles (one workgroup).
hipError_t launch_nrm2(hipStream_t s, int n, const double *x, double *r);
// Inside a launch group open on stream s, a norm of memory that an operation
// already recorded in the group writes must see the written values: it is
// kept and launched after the group's batched kernels (returns true).  Any
// other norm returns false and is launched at once, before them.
bool group_defer_nrm2(hipStream_t s, int n, const double *x, double *r);

// One CORE_dplgsy tile: the m x n block at global (m0, n0) of the gM x gM
// symmetric PLGSY matrix, into A (leading dimension lda).
struct PlgsyOp {
  double bump;
  double *A;
  long lda, gM, m0, n0;
  int m, n;
  unsigned long long seed;
};
// Batched PLGSY tiles (ops in device memory); max_n bounds every op's n.
hipError_t launch_plgsy_batched(hipStream_t s, const PlgsyOp *ops, int nops, int max_n);
// Inside a launch group open on stream s, records the tile for the group's
// batched launch (returns true); otherwise returns false.
bool group_record_plgsy(hipStream_t s, const PlgsyOp &op);

}  // namespace ttgc
