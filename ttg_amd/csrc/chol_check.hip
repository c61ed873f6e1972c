// The factorization check of examples/potrf/testing_dpotrf.cc:277-335 on
// the device, tile by tile:
//
//   result = ||L L^T - A0||_inf / (||A0||_inf N eps)
//
// * L L^T is formed in FP64 on MFMA from a row-packed copy P of the factor:
//   tile (m, k) of P sits at index m(m+1)/2 + k, so the tiles of one tile
//   row are consecutive and [L_m0 L_m1 ... L_mn] is one NB x (n+1)NB
//   column-major operand (ld = NB).  Output tile (m, n) is then ONE update
//   R_mn := A0_mn - [L_m0 .. L_mn][L_n0 .. L_nn]^T with k = (n+1) NB, and the
//   whole check is one batched launch.  P holds lower(L_kk) for diagonal
//   tiles (the trmm of :298-299 reads the lower triangle only) and zeros
//   outside the N x N matrix.
// * The residual is mirrored from its lower part (:302-310): row i sums the
//   lower entries of its row and, through the transpose, those of column i.
// * ||A0||_inf is lange(Inf) of A0 as getLAPACKMatrix returns it (pmw.h:
//   122-139; :312): the lower tiles as stored, diagonal tiles whole (their
//   upper part included), nothing above them.
//
// Row sums are reduced in a fixed order, so the result does not depend on
// the launch schedule.  The workspaces (P, R, partial sums) are library
// owned; the call waits for its kernels before it returns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <map>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

#include "chol_device.hpp"
#include "chol_internal.hpp"
#include "chol_kernels.hpp"
#include "chol_tiled.hpp"

namespace ttgc {

__device__ __forceinline__ void check_tile_mn(long t, int nt, int &m, int &n) {
  int lo = 0, hi = nt - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if ((long)mid * nt - (long)mid * (mid - 1) / 2 <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  n = lo;
  m = n + (int)(t - ((long)n * nt - (long)n * (n - 1) / 2));
}

// One workgroup per lower tile t = (m, n) of L: P tile (m, n) := L_mn with
// the strict upper part of a diagonal tile and every entry outside the
// N x N matrix set to zero.
__global__ __launch_bounds__(256) void rowpack_kernel(int N, int NB, int nt, const double *__restrict__ L,
                                                      double *__restrict__ P) {
  const long t = blockIdx.x;
  int m, n;
  check_tile_mn(t, nt, m, n);
  const long per = (long)NB * NB;
  const double *src = L + t * per;
  double *dst = P + ((long)m * (m + 1) / 2 + n) * per;
  const int mrows = min(NB, N - m * NB), ncols = min(NB, N - n * NB);
  for (long e = threadIdx.x; e < per; e += 256) {
    const int r = (int)(e % NB), c = (int)(e / NB);
    const bool keep = r < mrows && c < ncols && (m != n || r >= c);
    dst[e] = keep ? src[e] : 0.0;
  }
}

// R := A0 (all of the tile) and SA[t][r] = sum over the tile's row r of
// |A0| (the valid part; diagonal tiles whole).
__global__ __launch_bounds__(256) void check_prep_kernel(int N, int NB, int nt, const double *__restrict__ A0,
                                                         double *__restrict__ R, double *__restrict__ SA) {
  const long t = blockIdx.x;
  int m, n;
  check_tile_mn(t, nt, m, n);
  const long per = (long)NB * NB;
  const double *a = A0 + t * per;
  double *r = R + t * per;
  const int mrows = min(NB, N - m * NB), ncols = min(NB, N - n * NB);
  for (int i = threadIdx.x; i < NB; i += 256) {
    double s = 0.0;
    for (int c = 0; c < NB; ++c) {
      const double v = a[i + (long)c * NB];
      r[i + (long)c * NB] = v;
      if (i < mrows && c < ncols) s += fabs(v);
    }
    SA[t * NB + i] = s;
  }
}

// After the products: SR[t][0][r] = sum over row r of |R| (lower part of a
// diagonal tile), SR[t][1][c] = sum over column c of |R| (strictly lower
// part of a diagonal tile: the mirror's contribution to row n NB + c).
// Thread t owns rows t, t + 256 (NB <= 512); waves reduce a column by
// shuffles, and the four wave sums are added in wave order.
__global__ __launch_bounds__(256) void check_sums_kernel(int N, int NB, int nt, const double *__restrict__ R,
                                                         double *__restrict__ SR) {
  const long t = blockIdx.x;
  int m, n;
  check_tile_mn(t, nt, m, n);
  const double *a = R + t * (long)NB * NB;
  double *rs = SR + t * 2L * NB, *cs = rs + NB;
  __shared__ double csum[4][512];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int mrows = min(NB, N - m * NB), ncols = min(NB, N - n * NB);
  const bool diag = m == n;
  const int r0 = tid, r1 = tid + 256;
  double ra0 = 0.0, ra1 = 0.0;
  for (int c = 0; c < NB; ++c) {
    const long o = (long)c * NB;
    double u = 0.0;
    if (r0 < NB && r0 < mrows && c < ncols) {
      const double v = fabs(a[o + r0]);
      if (!diag || r0 >= c) ra0 += v;
      if (!diag || r0 > c) u += v;
    }
    if (r1 < NB && r1 < mrows && c < ncols) {
      const double v = fabs(a[o + r1]);
      if (!diag || r1 >= c) ra1 += v;
      if (!diag || r1 > c) u += v;
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) u += __shfl_xor(u, d);
    if (lane == 0) csum[wave][c] = u;
  }
  if (r0 < NB) rs[r0] = ra0;
  if (r1 < NB) rs[r1] = ra1;
  __syncthreads();
  for (int c = tid; c < NB; c += 256) cs[c] = ((csum[0][c] + csum[1][c]) + csum[2][c]) + csum[3][c];
}

// rows_r[i] = |R| row sum of the mirrored residual, rows_a[i] = |A0| row
// sum as stored; row block m: tiles (m, 0..m) by rows, then (m..nt-1, m) by
// columns.
__global__ void check_rows_kernel(int N, int NB, int nt, const double *__restrict__ SR, const double *__restrict__ SA,
                                  double *__restrict__ rows_r, double *__restrict__ rows_a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int m = i / NB, r = i % NB;
  double s = 0.0, sa = 0.0;
  for (int n = 0; n <= m; ++n) {
    const long t = tile_index(m, n, nt);
    s += SR[t * 2L * NB + r];
    sa += SA[t * NB + r];
  }
  for (int n2 = m; n2 < nt; ++n2) s += SR[tile_index(n2, m, nt) * 2L * NB + NB + r];
  rows_r[i] = s;
  rows_a[i] = sa;
}

// out = {||A0||_inf, ||L L^T - A0||_inf, ratio}; a NaN row sum of the
// residual propagates (testing_dpotrf.cc:319 fails on NaN).
__global__ __launch_bounds__(1024) void check_max_kernel(int N, const double *__restrict__ rows_r,
                                                         const double *__restrict__ rows_a, double *__restrict__ out) {
  __shared__ double sr[1024], sa[1024];
  double r = 0.0, a = 0.0;
  for (int i = threadIdx.x; i < N; i += 1024) {
    const double v = rows_r[i];
    r = (v > r || v != v) ? v : r;
    a = fmax(a, rows_a[i]);
  }
  sr[threadIdx.x] = r;
  sa[threadIdx.x] = a;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double o = sr[threadIdx.x + w];
      sr[threadIdx.x] = (o > sr[threadIdx.x] || o != o) ? o : sr[threadIdx.x];
      sa[threadIdx.x] = fmax(sa[threadIdx.x], sa[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sa[0];
    out[1] = sr[0];
    out[2] = sr[0] / (sa[0] * (double)N * DBL_EPSILON);
  }
}

template <typename T>
__global__ void widen_kernel(long n, const T *__restrict__ src, double *__restrict__ dst) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) dst[g] = (double)src[g];
}

template <typename T>
hipError_t launch_widen(hipStream_t s, long n, const T *src, double *dst) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(widen_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, src, dst);
  return hipGetLastError();
}
template hipError_t launch_widen<float>(hipStream_t, long, const float *, double *);
template hipError_t launch_widen<double>(hipStream_t, long, const double *, double *);

struct CheckWork {
  DevBuf P, R, sums, ops;
};
static std::mutex g_check_mu;
static std::unique_ptr<CheckWork> g_check;

void release_check_work() {
  std::lock_guard<std::mutex> lk(g_check_mu);
  if (!g_check) return;
  for (DevBuf *b : {&g_check->P, &g_check->R, &g_check->sums, &g_check->ops})
    if (b->ptr) (void)hipFree(b->ptr);
  g_check.reset();
}

}  // namespace ttgc

using namespace ttgc;

extern "C" int ttg_chol_dcheck_dpotrf_tiled(void *stream, int N, int NB, const double *A0, const double *L,
                                            double *d_out) {
  clear_error();
  if (N <= 0 || NB <= 0 || NB > TTG_CHOL_MAX_TILE || !A0 || !L || !d_out) {
    set_error("dcheck_dpotrf_tiled: bad argument (N=%d NB=%d)", N, NB);
    return TTG_CHOL_EINVAL;
  }
  hipStream_t s = as_stream(stream);
  const int nt = (N + NB - 1) / NB;
  const long ntiles = (long)nt * (nt + 1) / 2;
  const long per = (long)NB * NB;
  std::lock_guard<std::mutex> lk(g_check_mu);
  if (!g_check) g_check.reset(new CheckWork);
  CheckWork &w = *g_check;
  TTGC_HIP(w.P.reserve(sizeof(double) * (size_t)(ntiles * per)));
  TTGC_HIP(w.R.reserve(sizeof(double) * (size_t)(ntiles * per)));
  const size_t nsum = (size_t)ntiles * 3 * NB + 2 * (size_t)N;
  TTGC_HIP(w.sums.reserve(sizeof(double) * nsum));
  double *P = static_cast<double *>(w.P.ptr), *R = static_cast<double *>(w.R.ptr);
  double *SR = static_cast<double *>(w.sums.ptr), *SA = SR + ntiles * 2 * NB, *rows_r = SA + ntiles * NB;
  double *rows_a = rows_r + N;

  // one update per output tile, longest k first, dealt over the XCDs: the
  // update kernel hands XCD x the x-th eighth of the op list
  std::vector<std::pair<int, int>> mn;
  for (int n = nt - 1; n >= 0; --n)
    for (int m = n; m < nt; ++m) mn.emplace_back(m, n);
  const long len = ((long)mn.size() + kXcds - 1) / kXcds;
  std::vector<UpdOp<double>> ops((size_t)len * kXcds, UpdOp<double>{nullptr, nullptr, R, 0, 0, 0, NB, NB, NB, 0});
  bool fast = true;
  for (size_t i = 0; i < mn.size(); ++i) {
    const int m = mn[i].first, n = mn[i].second;
    UpdOp<double> op{P + (long)m * (m + 1) / 2 * per, P + (long)n * (n + 1) / 2 * per, R + tile_index(m, n, nt) * per,
                     NB, NB, (n + 1) * NB, NB, NB, NB, m == n ? 1 : 0};
    fast = fast && update_fast(op, NB);
    ops[(i % kXcds) * len + i / kXcds] = op;
  }
  TTGC_HIP(w.ops.reserve(sizeof(UpdOp<double>) * ops.size()));
  TTGC_HIP(hipMemcpyAsync(w.ops.ptr, ops.data(), sizeof(UpdOp<double>) * ops.size(), hipMemcpyHostToDevice, s));

  hipLaunchKernelGGL(rowpack_kernel, dim3((unsigned)ntiles), dim3(256), 0, s, N, NB, nt, L, P);
  TTGC_HIP(hipGetLastError());
  hipLaunchKernelGGL(check_prep_kernel, dim3((unsigned)ntiles), dim3(256), 0, s, N, NB, nt, A0, R, SA);
  TTGC_HIP(hipGetLastError());
  TTGC_HIP(launch_update<double>(s, static_cast<const UpdOp<double> *>(w.ops.ptr), (int)ops.size(), UpdOp<double>{},
                                 NB, NB, -1.0, 1.0, fast, nullptr, false));
  hipLaunchKernelGGL(check_sums_kernel, dim3((unsigned)ntiles), dim3(256), 0, s, N, NB, nt, R, SR);
  TTGC_HIP(hipGetLastError());
  hipLaunchKernelGGL(check_rows_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, N, NB, nt, SR, SA, rows_r,
                     rows_a);
  TTGC_HIP(hipGetLastError());
  hipLaunchKernelGGL(check_max_kernel, dim3(1), dim3(1024), 0, s, N, rows_r, rows_a, d_out);
  TTGC_HIP(hipGetLastError());
  // the workspaces and the host op list are reused by the next call
  TTGC_HIP(hipStreamSynchronize(s));
  return TTG_CHOL_OK;
}
