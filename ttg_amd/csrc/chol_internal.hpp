// Internal helpers shared by the C-ABI translation units of libttg_chol.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
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

}  // namespace ttgc
