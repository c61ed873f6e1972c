// Split-FP16 panel and trailing-update kernels (internal API).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ttgc {

// Split one FP32 panel tile (NB x NB, column-major, ld NB) into FP16 hi / lo
// parts (NB x NB each, k-slab-major: element (i, k) at (k / 32) NB 32 + 32 i
// + k % 32), each row i scaled by its own power of two s_i; rinv[i] = 1 / s_i
// (exact).
struct SplitOp {
  const float *src;
  uint16_t *hi, *lo;  // FP16 bit patterns
  float *rinv;        // NB row factors
};

// C (FP32 tile, column-major) -= diag(ra) (A_hi B_hi^T + A_hi B_lo^T + A_lo B_hi^T) diag(rb)
// for n x n tiles with inner dimension k; A/B parts are k-slab-major with ld
// rows per slab; ra / rb are the row factors of the A and B splits.  With
// npan > 1, the products of npan consecutive panels are subtracted one
// after the other in the same pass over C, each rounded to FP32 as separate
// updates would be; panel j's parts sit pstride halves and its row factors
// sstride floats after panel 0's.
struct F16UpdOp {
  const uint16_t *ahi, *alo, *bhi, *blo;
  const float *ra, *rb;
  float *C;
  int n, k, ld, ldc;
  int lower;
  int npan;  // 0 or 1: one product
  long pstride, sstride;
};

// low-precision trailing-update modes
constexpr int kLowF16x2 = 1;  // FP16 hi + lo split, three MFMA products (~FP32 accuracy)
constexpr int kLowF16 = 2;    // one FP16 product (operands rounded to FP16 after row scaling)
constexpr int kLowBF16 = 3;   // one BF16 product

hipError_t launch_split(hipStream_t s, const SplitOp *ops, int nops, int NB, int mode);
hipError_t launch_update_f16x2(hipStream_t s, const F16UpdOp *ops, int nops, int NB, const int *d_abort, int mode);

}  // namespace ttgc
