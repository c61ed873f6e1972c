// Host-visible declarations of the tile-operation descriptors and the kernel
// launchers of libttg_chol (internal C++ API, not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ttgc {

// One POTRF of an n x n tile (n <= 512).
struct PotrfOp {
  double *A;
  double *dinv;  // optional: receives inv(L_jj) of every 32 x 32 diagonal block, 32x32 col-major each
  int *info;     // optional device int
  int n, lda;
  int info_offset;
  int write_zero;  // 1: store 0 into *info on success (per-tile API semantics)
};

// One in-place TRSM  B := B * L^{-T}  of an m x n tile (n <= 512).
template <typename T>
struct TrsmOp {
  const T *L;
  T *B;
  const T *dinv;  // inv(L_jj) blocks as produced by the POTRF kernel (or the dinv kernel)
  int m, n, ldl, ldb;
};

// One update  C := alpha A B^T + beta C  (lower triangle only if lower != 0).
template <typename T>
struct UpdOp {
  const T *A;
  const T *B;
  T *C;
  int m, n, k;
  int lda, ldb, ldc;
  int lower;
};

// Launchers.  `ops` is a DEVICE array of nops descriptors, or nullptr, in
// which case the single descriptor `one` (passed by value as a kernel
// argument) is used and nops must be 1.  `d_abort` (optional device int):
// when nonzero at kernel start, the kernel does nothing (a failed POTRF
// stops the rest of the factorization without a host round trip).
hipError_t launch_potrf(hipStream_t s, const PotrfOp *ops, int nops, const PotrfOp &one, const int *d_abort);
// inverses of the 32 x 32 diagonal blocks of an already factored tile
hipError_t launch_dinv(hipStream_t s, const PotrfOp *ops, int nops, const PotrfOp &one);
template <typename T>
hipError_t launch_trsm(hipStream_t s, const TrsmOp<T> *ops, int nops, const TrsmOp<T> &one, int max_m,
                       const int *d_abort);
// fast: every op has m % 128 == 0, n % 128 == 0, k % BK == 0, leading
// dimensions a multiple of the 16-byte vector and 16-byte aligned pointers.
template <typename T>
hipError_t launch_update(hipStream_t s, const UpdOp<T> *ops, int nops, const UpdOp<T> &one, int max_m, int max_n,
                         T alpha, T beta, bool fast, const int *d_abort);
// K granularity of the fast update path
template <typename T>
constexpr int update_bk() { return sizeof(T) == 8 ? 16 : 32; }

}  // namespace ttgc
