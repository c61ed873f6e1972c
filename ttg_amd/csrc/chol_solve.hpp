// Triangular-solve plans and residuals on the tiled factor (internal API).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "chol_internal.hpp"
#include "chol_kernels.hpp"
#include "chol_sweep.hpp"

namespace ttgc {

// Forward + backward sweeps for nrhs (<= 32) right-hand sides held
// transposed in d_bt (nrhs x Np, Np = NT*NB, zero padded).  T is the
// element type of the factor tiles.  prepare() inverts every diagonal tile
// of the factor (d_linv); the sweeps are then skinny GEMMs only:
//   forward   Xt_i = Bt_i Linv_i^T,  Bt_m -= Xt_i L_mi^T  (m > i)
//   backward  Bt_i = Xt_i Linv_i,    Xt_m -= Bt_i L_im    (m < i)
// leaving the solution in d_bt.
template <typename T>
struct SolvePlan {
  int N = 0, NB = 0, NT = 0, Np = 0, nrhs = 0, dstride = 0;
  const T *tiles = nullptr;
  DevBuf d_ops, d_bt, d_xt, d_dinv, d_linv, d_part;  // d_part: k-split chunks of one inverse apply
  size_t o_dj = 0, o_inv = 0, o_upd = 0;
  std::vector<int> fw_app, fw_upd_off, fw_upd_cnt, bw_app, bw_upd_off, bw_upd_cnt;
  std::vector<UpdOp<T>> host_ops;  // host copy of the sweep descriptors (the applies read theirs from it)
  // persistent sweeps (the SWEEP knob > 0): both sweeps as one launch of
  // sweep_grid workgroups per 16 right-hand sides, counters in d_ctr
  bool persist = false;
  int sweep_grid = 0;
  long long sweep_timeout = 0;
  DevBuf d_ctr;
  DevBuf d_pr;  // the folded products P_{m,j}, then R_{i,j} (kSweepDepth (NT-1) tiles each), from prepare()
  size_t o_pr = 0;  // offset of their descriptors in d_ops
  int n_pr = 0;     // P descriptors (the R descriptors follow)
  size_t n_pr_all = 0;
  hipStream_t s_main = nullptr;  // the stream the sweep graph is captured on
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;

  ~SolvePlan();
  int build();
  int prepare(hipStream_t s);  // after every (re)factorization of the tiles
  int enqueue_sweeps(hipStream_t s);
  int run_sweeps(hipStream_t user);
  PotrfOpT<T> *dj_ops() { return reinterpret_cast<PotrfOpT<T> *>(static_cast<char *>(d_ops.ptr) + o_dj); }
  TrsmOp<T> *inv_ops() { return reinterpret_cast<TrsmOp<T> *>(static_cast<char *>(d_ops.ptr) + o_inv); }
  UpdOp<T> *upd_ops() { return reinterpret_cast<UpdOp<T> *>(static_cast<char *>(d_ops.ptr) + o_upd); }
  UpdOp<T> *pr_ops() { return reinterpret_cast<UpdOp<T> *>(static_cast<char *>(d_ops.ptr) + o_pr); }
  T *bt() { return static_cast<T *>(d_bt.ptr); }
  // The failure word of the persistent sweeps (nullptr when the plan does
  // not use them): nonzero after a solve in which a launch gave up (time
  // limit; the solve's results are then not valid); the next solve clears it.
  const int *fail_word() const {
    return persist ? static_cast<const int *>(d_ctr.ptr) + kSweepFail * kSweepStride : nullptr;
  }
};

template <typename T>
SolvePlan<T> *get_solve_plan(int N, int NB, const T *tiles, int nrhs, int *rc);

template <typename TS, typename TD>
hipError_t launch_to_rows(hipStream_t s, int N, int Np, int nrhs, const TS *B, long ldb, TD *Bt);
template <typename TS, typename TD>
hipError_t launch_from_rows(hipStream_t s, int N, int nrhs, const TS *Bt, TD *B, long ldb);

// R := B - A X
int residual_fp64(hipStream_t s, int N, int NB, const double *tiles, int nrhs, const double *X, long ldx,
                  const double *B, long ldb, double *R, long ldr);

}  // namespace ttgc
