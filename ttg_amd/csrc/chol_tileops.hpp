// Tile passes of the per-tile POINV operations (TRTRI, LAUUM, the left TRSM
// and the TRMM of examples/potrf/trtri_L.h and lauum.h): copies,
// transposes and triangle extractions of one tile into or out of a work
// tile, batched over a launch group like the other per-tile kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace ttgc {

enum TileOpMode : int {
  kTopCopy = 0,        // dst (m x n) := alpha src (may be in place)
  kTopTrans = 1,       // dst (n x m) := alpha src^T
  kTopLowerZero = 2,   // dst (n x n) := lower(src), strict upper set to zero
  kTopTransLower = 3,  // lower part of dst (n x n) := lower part of src^T (dst's strict upper untouched)
  kTopIdentity = 4,    // dst (n x n) := I
};

struct TopOp {
  const double *src;
  double *dst;
  int m, n;  // of src (kTopIdentity: of dst)
  int lds, ldd;
  int mode;
  double alpha;
};

// *info := info_offset + 1-based index of the first exact zero on the
// diagonal of the n x n tile A (LAPACK dtrtri's singularity test), or 0
struct DiagZeroOp {
  const double *A;
  int *info;
  int n, lda, info_offset;
};

hipError_t launch_tile_ops(hipStream_t s, const TopOp *ops, int nops, int max_m, int max_n);
hipError_t launch_diag_zero(hipStream_t s, const DiagZeroOp *ops, int nops);

}  // namespace ttgc
