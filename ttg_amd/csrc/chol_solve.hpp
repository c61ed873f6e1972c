// Triangular-solve plans and residuals on the tiled factor (internal API).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "chol_internal.hpp"
#include "chol_kernels.hpp"

namespace ttgc {

// Forward + backward sweeps for nrhs (<= 32) right-hand sides held
// transposed in d_bt (nrhs x Np, Np = NT*NB, zero padded).  T is the
// element type of the factor tiles.
template <typename T>
struct SolvePlan {
  int N = 0, NB = 0, NT = 0, Np = 0, nrhs = 0, dstride = 0;
  const T *tiles = nullptr;
  DevBuf d_ops, d_bt, d_dinv;
  size_t o_dj = 0, o_trsm = 0, o_upd = 0;
  std::vector<int> fw_trsm, fw_upd_off, fw_upd_cnt, bw_trsm, bw_upd_off, bw_upd_cnt;
  hipStream_t s_main = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;

  ~SolvePlan();
  int build();
  int enqueue_sweeps(hipStream_t s);
  int run_sweeps(hipStream_t user);
  PotrfOpT<T> *dj_ops() { return reinterpret_cast<PotrfOpT<T> *>(static_cast<char *>(d_ops.ptr) + o_dj); }
  TrsmOp<T> *trsm_ops() { return reinterpret_cast<TrsmOp<T> *>(static_cast<char *>(d_ops.ptr) + o_trsm); }
  UpdOp<T> *upd_ops() { return reinterpret_cast<UpdOp<T> *>(static_cast<char *>(d_ops.ptr) + o_upd); }
  T *bt() { return static_cast<T *>(d_bt.ptr); }
};

template <typename T>
SolvePlan<T> *get_solve_plan(int N, int NB, const T *tiles, int nrhs, int *rc);

template <typename TS, typename TD>
hipError_t launch_to_rows(hipStream_t s, int N, int Np, int nrhs, const TS *B, long ldb, TD *Bt);
template <typename TS, typename TD>
hipError_t launch_from_rows(hipStream_t s, int N, int nrhs, const TS *Bt, TD *B, long ldb);

// R := B - A X (abs_a: B - |A| X, for row sums of |A|)
int residual_fp64(hipStream_t s, int N, int NB, const double *tiles, int nrhs, const double *X, long ldx,
                  const double *B, long ldb, double *R, long ldr, bool abs_a = false);

}  // namespace ttgc
