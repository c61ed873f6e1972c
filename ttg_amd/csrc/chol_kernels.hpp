// Host-visible declarations of the tile-operation descriptors and the kernel
// launchers of libttg_chol (internal C++ API, not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ttgc {

// One POTRF of an n x n tile (n <= 512) stored as TS (double, or float for
// the mixed-precision factor); the arithmetic is FP64 either way.
template <typename TS>
struct PotrfOpT {
  TS *A;
  TS *dinv;   // optional: receives inv(L_jj) of every 32 x 32 diagonal block, 32x32 col-major each
  int *info;  // optional device int
  int n, lda;
  int info_offset;
  int write_zero;  // 1: store 0 into *info on success (per-tile API semantics)
  double *pw;      // FP64 workspace (n x n) for float tiles; unused for double tiles
};
using PotrfOp = PotrfOpT<double>;

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
// POTRF computes in FP64 (float tiles through the FP64 workspace pw), or,
// for float tiles with compute_ts, in FP32 on the tile itself (pw unused):
// for factorizations whose trailing updates carry larger errors anyway.
template <typename TS>
hipError_t launch_potrf(hipStream_t s, const PotrfOpT<TS> *ops, int nops, const PotrfOpT<TS> &one,
                        const int *d_abort, bool compute_ts = false);
// inverses of the 32 x 32 diagonal blocks of an already factored tile
template <typename TS>
hipError_t launch_dinv(hipStream_t s, const PotrfOpT<TS> *ops, int nops, const PotrfOpT<TS> &one);
// trans: B := B L^{-T}; otherwise B := B L^{-1}
template <typename T>
hipError_t launch_trsm(hipStream_t s, const TrsmOp<T> *ops, int nops, const TrsmOp<T> &one, int max_m,
                       const int *d_abort, bool trans = true);
// fast: every op satisfies update_fast() (see below).  Outputs with at most
// 32 rows (max_m <= 32) use a 32 x 64 block shape.
// tb: the B operand's element (j, k) is read from B[k + j*ldb], i.e. the
// update is C := alpha A B + beta C with B stored k x n column-major.
// ta (FP64 only): the A operand's element (i, k) is read from A[k + i*lda],
// i.e. op(A) = A^T with A stored k x m column-major.
template <typename T>
hipError_t launch_update(hipStream_t s, const UpdOp<T> *ops, int nops, const UpdOp<T> &one, int max_m, int max_n,
                         T alpha, T beta, bool fast, const int *d_abort, bool tb = false, bool ta = false);
// C := alpha A B^T + beta C for the SPD generator (same code as the update,
// separate kernel symbol).
hipError_t launch_gen_gemm(hipStream_t s, const UpdOp<double> *ops, int nops, int max_m, int max_n, double alpha,
                           double beta, bool fast);

// Whether an update op may take the unpredicated vector path: full 128-wide
// column blocks, k a multiple of 32, 16-byte aligned operands; plus full
// 128-row blocks unless the op is skinny (m <= 32, rows always predicated).
template <typename T>
inline bool update_fast(const UpdOp<T> &op, int max_m) {
  constexpr int V = 16 / sizeof(T);
  auto al = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  bool ok = op.n % 128 == 0 && op.k % 32 == 0 && op.ldb % V == 0 && al(op.B);
  if (max_m > 32) ok = ok && op.m % 128 == 0 && op.lda % V == 0 && op.ldc % V == 0 && al(op.A) && al(op.C);
  return ok;
}

}  // namespace ttgc
