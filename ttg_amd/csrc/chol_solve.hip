// Solves and residuals on the tiled factor (no reference counterpart; the
// building blocks of iterative refinement).
//
//   potrs   : B := (L L^T)^{-1} B.  The right-hand sides are held transposed
//             (Bt = B^T, nrhs x N).  The diagonal tiles are inverted once per
//             factor (batched TRSM of identity tiles, all tiles in one launch),
//             so both sweeps are skinny GEMMs with a fixed, short latency per
//             step instead of a sequential triangular solve:
//             forward  Xt_i = Bt_i Linv_i^T, Bt_m -= Xt_i L_mi^T (m > i);
//             backward Bt_i = Xt_i Linv_i,   Xt_m -= Bt_i L_im   (m < i).
//             Every step is two launches; the sequence is captured in a
//             hipGraph.
//   residual: R := B - A X with A symmetric from its lower tiles.  Each tile
//             (m, n) is read once; a workgroup forms A_mn X_n and A_mn^T X_m
//             on FP64 MFMA from 64 x 64 LDS blocks and stores them to per-tile
//             slots, and a second kernel sums the slots in a fixed order
//             (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#include "chol_device.hpp"
#include "chol_internal.hpp"
#include "chol_kernels.hpp"
#include "chol_solve.hpp"
#include "chol_sweep.hpp"
#include "chol_tiled.hpp"

namespace ttgc {

constexpr int kRhsChunk = 16;

// ---------------------------------------------------------------------------
// transposition of the right-hand sides (zero padding beyond N)
// ---------------------------------------------------------------------------
template <typename TS, typename TD>
__global__ void to_rows_kernel(int N, int Np, int nrhs, const TS *__restrict__ B, long ldb, TD *__restrict__ Bt) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)Np * nrhs) return;
  const int j = (int)(g % nrhs), i = (int)(g / nrhs);
  Bt[g] = (i < N) ? (TD)B[i + (long)j * ldb] : TD(0);
}

template <typename TS, typename TD>
__global__ void from_rows_kernel(int N, int nrhs, const TS *__restrict__ Bt, TD *__restrict__ B, long ldb) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)N * nrhs) return;
  const int i = (int)(g % N), j = (int)(g / N);
  B[i + (long)j * ldb] = (TD)Bt[j + (long)i * nrhs];
}

template <typename TS, typename TD>
hipError_t launch_to_rows(hipStream_t s, int N, int Np, int nrhs, const TS *B, long ldb, TD *Bt) {
  const long n = (long)Np * nrhs;
  hipLaunchKernelGGL((to_rows_kernel<TS, TD>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, N, Np, nrhs, B,
                     ldb, Bt);
  return hipGetLastError();
}
template <typename TS, typename TD>
hipError_t launch_from_rows(hipStream_t s, int N, int nrhs, const TS *Bt, TD *B, long ldb) {
  const long n = (long)N * nrhs;
  hipLaunchKernelGGL((from_rows_kernel<TS, TD>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, N, nrhs, Bt, B,
                     ldb);
  return hipGetLastError();
}

template hipError_t launch_to_rows<double, double>(hipStream_t, int, int, int, const double *, long, double *);
template hipError_t launch_from_rows<double, double>(hipStream_t, int, int, const double *, double *, long);
template hipError_t launch_to_rows<double, float>(hipStream_t, int, int, int, const double *, long, float *);
template hipError_t launch_from_rows<float, double>(hipStream_t, int, int, const float *, double *, long);

// ---------------------------------------------------------------------------
// triangular-solve plan
// ---------------------------------------------------------------------------
// identity into every diagonal-tile slot of the inverse (dim x dim, zero padding)
template <typename T>
__global__ void identity_tiles_kernel(int N, int NB, T *__restrict__ linv) {
  const int k = blockIdx.y;
  const int dim = min(NB, N - k * NB);
  T *t = linv + (long)k * NB * NB;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < (long)NB * NB; e += (long)gridDim.x * blockDim.x) {
    const int r = (int)(e % NB), c = (int)(e / NB);
    t[e] = (r == c && r < dim) ? T(1) : T(0);
  }
}

// ---------------------------------------------------------------------------
// The inverse apply of a sweep step, split over k:
//   P[y][c][r] = sum_{kk in chunk y} A(r, kk) B(c, kk),   A (nr x k), B (n x k),
// B(c, kk) = B[c + kk ldb] (TB: B[kk + c ldb]).  Workgroup (x, y) takes the
// 64 columns x and the 64-k chunk y, so a 512 x 512 tile is streamed by 64
// workgroups with all their loads in flight at once.  apply_reduce_kernel
// adds the chunks in order (deterministic).
// ---------------------------------------------------------------------------
constexpr int kApB = 64;
constexpr int kMaxChunks = 8;  // k chunks of a product: NB <= 512

template <typename T, bool TB>
__global__ __launch_bounds__(256) void apply_partial_kernel(int nr, int n, int k, const T *__restrict__ A, int lda,
                                                            const T *__restrict__ B, int ldb, T *__restrict__ P) {
  const int c0 = blockIdx.x * kApB, k0 = blockIdx.y * kApB;
  __shared__ T bs[kApB][kApB + 1];  // bs[kk][cc] = B(c0 + cc, k0 + kk)
  __shared__ T as[kApB][17];        // as[kk][r] = A(r0 + r, k0 + kk), 16 right-hand sides at a time
  const int tid = threadIdx.x, cc = tid & 63, rg = tid >> 6;
  T bv[kApB * kApB / 256];
#pragma unroll
  for (int i = 0; i < kApB * kApB / 256; ++i) {
    const int e = tid + 256 * i, fast = e & 63, slow = e >> 6;  // fast runs along contiguous memory
    const int c = TB ? slow : fast, kk = TB ? fast : slow;
    const bool ok = c0 + c < n && k0 + kk < k;
    bv[i] = ok ? (TB ? B[(k0 + kk) + (long)(c0 + c) * ldb] : B[(c0 + c) + (long)(k0 + kk) * ldb]) : T(0);
  }
#pragma unroll
  for (int i = 0; i < kApB * kApB / 256; ++i) {
    const int e = tid + 256 * i, fast = e & 63, slow = e >> 6;
    bs[TB ? fast : slow][TB ? slow : fast] = bv[i];
  }
  for (int r0 = 0; r0 < nr; r0 += 16) {
    __syncthreads();
    T av[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = tid + 256 * i, r = e & 15, kk = e >> 4;
      const bool ok = r0 + r < nr && k0 + kk < k;
      av[i] = ok ? A[(r0 + r) + (long)(k0 + kk) * lda] : T(0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = tid + 256 * i;
      as[e >> 4][e & 15] = av[i];
    }
    __syncthreads();
    T acc[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll 8
    for (int kk = 0; kk < kApB; ++kk) {
      const T bb = bs[kk][cc];
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] += as[kk][rg * 4 + j] * bb;
    }
    if (c0 + cc < n) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = r0 + rg * 4 + j;
        if (r < nr) P[((long)blockIdx.y * n + c0 + cc) * nr + r] = acc[j];
      }
    }
  }
}

template <typename T>
__global__ void apply_reduce_kernel(int nr, int n, int nkc, const T *__restrict__ P, T *__restrict__ C, int ldc) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nr * n) return;
  const int r = g % nr, c = g / nr;
  T pv[kMaxChunks];
#pragma unroll
  for (int q = 0; q < kMaxChunks; ++q) pv[q] = q < nkc ? P[((long)q * n + c) * nr + r] : T(0);
  T s = T(0);
#pragma unroll
  for (int q = 0; q < kMaxChunks; ++q) s += pv[q];
  C[(long)c * ldc + r] = s;
}

// P := the k-split product of op (A B^T).
template <typename T>
static hipError_t launch_partial(hipStream_t s, const UpdOp<T> &op, bool tb, T *P) {
  const dim3 g((op.n + kApB - 1) / kApB, (op.k + kApB - 1) / kApB);
  if (tb)
    hipLaunchKernelGGL((apply_partial_kernel<T, true>), g, dim3(256), 0, s, op.m, op.n, op.k, op.A, op.lda, op.B,
                       op.ldb, P);
  else
    hipLaunchKernelGGL((apply_partial_kernel<T, false>), g, dim3(256), 0, s, op.m, op.n, op.k, op.A, op.lda, op.B,
                       op.ldb, P);
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_reduce(hipStream_t s, const UpdOp<T> &op, const T *P) {
  const int tot = op.m * op.n;
  hipLaunchKernelGGL(apply_reduce_kernel<T>, dim3((tot + 255) / 256), dim3(256), 0, s, op.m, op.n,
                     (op.k + kApB - 1) / kApB, P, op.C, op.ldc);
  return hipGetLastError();
}

template <typename T>
int SolvePlan<T>::build() {
  const int wgs = plan_knob(TTG_CHOL_KNOB_SWEEP, "TTG_CHOL_SWEEP", 2);
  persist = wgs > 0 && NB <= kSweepMaxNB && NB % kSweepRows == 0;
  NT = (N + NB - 1) / NB;
  Np = NT * NB;
  dstride = ((NB + 31) / 32) * 32 * 32;
  TTGC_HIP(d_bt.reserve(sizeof(T) * (size_t)Np * nrhs));
  TTGC_HIP(d_xt.reserve(sizeof(T) * (size_t)Np * nrhs));
  TTGC_HIP(d_dinv.reserve(sizeof(T) * (size_t)dstride * NT));
  TTGC_HIP(d_linv.reserve(sizeof(T) * (size_t)NB * NB * NT));
  // the inverse apply's k chunks
  TTGC_HIP(d_part.reserve(sizeof(T) * (size_t)((NB + kApB - 1) / kApB) * NB * nrhs));
  T *bt = static_cast<T *>(d_bt.ptr);
  T *xt = static_cast<T *>(d_xt.ptr);
  T *dinv = static_cast<T *>(d_dinv.ptr);
  T *linv = static_cast<T *>(d_linv.ptr);
  std::vector<PotrfOpT<T>> dj;
  std::vector<TrsmOp<T>> inv;
  std::vector<UpdOp<T>> upd;
  auto tile = [&](int m, int n) { return tiles + tile_index(m, n, NT) * (long)NB * NB; };
  auto li = [&](int i) { return linv + (long)i * NB * NB; };
  auto rows = [&](int i) { return bt + (long)i * NB * nrhs; };
  auto xrows = [&](int i) { return xt + (long)i * NB * nrhs; };
  // actual (edge) sizes everywhere: the solve never reads tile padding
  auto dim = [&](int i) { return std::min(NB, N - i * NB); };
  for (int k = 0; k < NT; ++k) {
    dj.push_back(PotrfOpT<T>{const_cast<T *>(tile(k, k)), dinv + (long)k * dstride, nullptr, dim(k), NB, 0, 0});
    // Linv_k := I L_kk^{-1}
    inv.push_back(TrsmOp<T>{tile(k, k), li(k), dinv + (long)k * dstride, dim(k), dim(k), NB, NB});
  }
  // forward: apply (beta 0, B^T), then the updates (B^T)
  for (int i = 0; i < NT; ++i) {
    fw_app.push_back((int)upd.size());
    upd.push_back(UpdOp<T>{rows(i), li(i), xrows(i), nrhs, dim(i), dim(i), nrhs, NB, nrhs, 0});
    fw_upd_off.push_back((int)upd.size());
    for (int m = i + 1; m < NT; ++m)
      upd.push_back(UpdOp<T>{xrows(i), tile(m, i), rows(m), nrhs, dim(m), dim(i), nrhs, NB, nrhs, 0});
    fw_upd_cnt.push_back((int)upd.size() - fw_upd_off.back());
  }
  // backward: apply (beta 0, B as stored), then the updates (B as stored)
  for (int i = NT - 1; i >= 0; --i) {
    bw_app.push_back((int)upd.size());
    upd.push_back(UpdOp<T>{xrows(i), li(i), rows(i), nrhs, dim(i), dim(i), nrhs, NB, nrhs, 0});
    bw_upd_off.push_back((int)upd.size());
    for (int m = 0; m < i; ++m)
      upd.push_back(UpdOp<T>{rows(i), tile(i, m), xrows(m), nrhs, dim(m), dim(i), nrhs, NB, nrhs, 0});
    bw_upd_cnt.push_back((int)upd.size() - bw_upd_off.back());
  }
  // the persistent sweeps' folded off-diagonal products (prepare() forms
  // them): P_{m,j} = Linv_m L_{m,m-j} and R_{i,j} = L_{i+j,i} Linv_i
  std::vector<UpdOp<T>> pr;
  if (persist && NT > 1) {
    const long per = (long)(NT - 1) * NB * NB;  // tiles of one j and direction
    TTGC_HIP(d_pr.reserve(sizeof(T) * (size_t)2 * kSweepDepth * per));
    T *pbase = static_cast<T *>(d_pr.ptr), *rbase = pbase + kSweepDepth * per;
    for (int j = 1; j <= kSweepDepth; ++j)
      for (int m = j; m < NT; ++m)  // C = Linv_m B^T, B(c, k) = L_{m,m-j}(k, c)
        pr.push_back(UpdOp<T>{li(m), tile(m, m - j), pbase + (j - 1) * per + (long)(m - j) * NB * NB, dim(m),
                              dim(m - j), dim(m), NB, NB, NB, 0});
    n_pr = (int)pr.size();
    for (int j = 1; j <= kSweepDepth; ++j)
      for (int i = 0; i + j < NT; ++i)  // C = L_{i+j,i} B^T, B(c, k) = Linv_i(k, c)
        pr.push_back(UpdOp<T>{tile(i + j, i), li(i), rbase + (j - 1) * per + (long)i * NB * NB, dim(i + j), dim(i),
                              dim(i), NB, NB, NB, 0});
  }
  auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
  o_dj = 0;
  o_inv = align(dj.size() * sizeof(PotrfOpT<T>));
  o_upd = align(o_inv + inv.size() * sizeof(TrsmOp<T>));
  o_pr = align(o_upd + upd.size() * sizeof(UpdOp<T>));
  const size_t total = align(o_pr + pr.size() * sizeof(UpdOp<T>));
  TTGC_HIP(d_ops.reserve(total));
  std::vector<char> host(total);
  memcpy(host.data() + o_dj, dj.data(), dj.size() * sizeof(PotrfOpT<T>));
  memcpy(host.data() + o_inv, inv.data(), inv.size() * sizeof(TrsmOp<T>));
  if (!upd.empty()) memcpy(host.data() + o_upd, upd.data(), upd.size() * sizeof(UpdOp<T>));
  if (!pr.empty()) memcpy(host.data() + o_pr, pr.data(), pr.size() * sizeof(UpdOp<T>));
  n_pr_all = pr.size();
  TTGC_HIP(hipMemcpy(d_ops.ptr, host.data(), total, hipMemcpyHostToDevice));
  host_ops.assign(upd.begin(), upd.end());
  TTGC_HIP(hipStreamCreateWithFlags(&s_main, hipStreamNonBlocking));
  if (persist) {
    int dev = 0, ncu = 0, per_cu = 0, khz = 0;
    TTGC_HIP(hipGetDevice(&dev));
    TTGC_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    TTGC_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    TTGC_HIP(sweep_occupancy<T>(&per_cu));
    // every workgroup resident at once (and at least NB/32 of them: the
    // sweep's progress condition)
    sweep_grid = std::min(wgs, per_cu) * ncu;
    if (sweep_grid < NB / kSweepRows) {
      set_error("solve plan: %d resident sweep workgroups, %d needed", sweep_grid, NB / kSweepRows);
      return TTG_CHOL_ENOTSUP;
    }
    sweep_timeout = (long long)(khz > 0 ? khz : 100000) * 1000LL * 10;  // 10 s
    const size_t nc = (size_t)sweep_counter_ints(NT);
    TTGC_HIP(d_ctr.reserve(sizeof(int) * nc));
    TTGC_HIP(hipMemset(d_ctr.ptr, 0, sizeof(int) * nc));
  }
  return TTG_CHOL_OK;
}

template <typename T>
int SolvePlan<T>::prepare(hipStream_t s) {
  PotrfOpT<T> pnone{};
  TrsmOp<T> tnone{};
  TTGC_HIP(launch_dinv(s, dj_ops(), NT, pnone));
  hipLaunchKernelGGL(identity_tiles_kernel<T>, dim3(64, NT), dim3(256), 0, s, N, NB, static_cast<T *>(d_linv.ptr));
  TTGC_HIP(hipGetLastError());
  TTGC_HIP(launch_trsm<T>(s, inv_ops(), NT, tnone, NB, nullptr, false));
  if (persist && n_pr_all > 0) {
    UpdOp<T> unone{};
    const bool fast = NB % 128 == 0 && N % NB == 0;
    TTGC_HIP(launch_update<T>(s, pr_ops(), (int)n_pr_all, unone, NB, NB, T(1), T(0), fast, nullptr, true));
  }
  return TTG_CHOL_OK;
}

// Sweep step q solves block i (x_i known) and prepares block i' = i +- 1,
// three launches on one stream: (1) the updates of every remaining block
// from x_i, one batched launch (HBM-bound early on: the whole column of L
// tiles is read), (2) the inverse apply of block i' as a k-split product
// (every load in flight at once), (3) its chunk reduction, giving x_i'.  The
// sweeps are replayed as one graph on one stream: kernels then follow each
// other with no gap, while a dependency across queues costs ~20 us on this
// GPU and a directly launched step is bound by the host's launch calls.
template <typename T>
int SolvePlan<T>::enqueue_sweeps(hipStream_t s) {
  if (persist) {
    // passes of up to 16 columns: a column's arithmetic does not depend on
    // how many columns the solve has
    for (int j0 = 0; j0 < nrhs; j0 += kSweepMaxRhs) {
      const T *pm = static_cast<const T *>(d_pr.ptr);
      SweepArgs<T> a{tiles,
                     static_cast<const T *>(d_linv.ptr),
                     pm,
                     pm + (long)kSweepDepth * (NT - 1) * NB * NB,
                     bt() + j0,
                     static_cast<T *>(d_xt.ptr) + j0,
                     static_cast<int *>(d_ctr.ptr),
                     N,
                     NB,
                     NT,
                     std::min(kSweepMaxRhs, nrhs - j0),
                     nrhs,
                     sweep_full<T>(N, NB) ? 1 : 0,
                     sweep_timeout};
      TTGC_HIP(launch_sweep<T>(s, a, sweep_grid, j0 == 0));
    }
    return TTG_CHOL_OK;
  }
  UpdOp<T> unone{};
  const T m1 = T(-1), p1 = T(1);
  const bool fast = NB % 128 == 0 && N % NB == 0;
  T *pa = static_cast<T *>(d_part.ptr);  // the apply's chunks
  for (int dir = 0; dir < 2; ++dir) {
    const bool tb = dir == 1;
    const std::vector<int> &app = tb ? bw_app : fw_app, &off = tb ? bw_upd_off : fw_upd_off,
                           &cnt = tb ? bw_upd_cnt : fw_upd_cnt;
    TTGC_HIP(launch_partial<T>(s, host_ops[app[0]], tb, pa));
    TTGC_HIP(launch_reduce<T>(s, host_ops[app[0]], pa));
    for (int q = 0; q < NT; ++q) {
      const int c = cnt[q];
      if (c == 0) break;
      TTGC_HIP(launch_update<T>(s, upd_ops() + off[q], c, unone, nrhs, NB, m1, p1, fast, nullptr, tb));
      TTGC_HIP(launch_partial<T>(s, host_ops[app[q + 1]], tb, pa));
      TTGC_HIP(launch_reduce<T>(s, host_ops[app[q + 1]], pa));
    }
  }
  return TTG_CHOL_OK;
}

template <typename T>
SolvePlan<T>::~SolvePlan() {
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  if (s_main) (void)hipStreamDestroy(s_main);
  for (DevBuf *b : {&d_ops, &d_bt, &d_xt, &d_dinv, &d_linv, &d_part, &d_ctr, &d_pr})
    if (b->ptr) (void)hipFree(b->ptr);
}

// Run the two sweeps on Bt (already filled) as a replayed graph (captured on
// the first run), ordered after `user` and before whatever `user` does next.
template <typename T>
int SolvePlan<T>::run_sweeps(hipStream_t user) {
  if (!gexec) {
    TTGC_HIP(hipStreamBeginCapture(s_main, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_sweeps(s_main);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s_main, &g);
    if (rc != TTG_CHOL_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    TTGC_HIP(e);
    graph = g;
    TTGC_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
  }
  TTGC_HIP(hipGraphLaunch(gexec, user));
  return TTG_CHOL_OK;
}

template struct SolvePlan<double>;
template struct SolvePlan<float>;

static std::mutex g_solve_mu;
static std::map<std::tuple<int, int, const void *, int, int>, std::unique_ptr<SolvePlan<double>>> g_dplans;
static std::map<std::tuple<int, int, const void *, int, int>, std::unique_ptr<SolvePlan<float>>> g_splans;

template <typename T>
SolvePlan<T> *get_solve_plan(int N, int NB, const T *tiles, int nrhs, int *rc) {
  std::lock_guard<std::mutex> lk(g_solve_mu);
  auto key = std::make_tuple(N, NB, (const void *)tiles, nrhs, (int)sizeof(T));
  auto &plans = [&]() -> auto & {
    if constexpr (sizeof(T) == 8)
      return g_dplans;
    else
      return g_splans;
  }();
  auto &p = plans[key];
  if (!p) {
    p.reset(new SolvePlan<T>);
    p->N = N;
    p->NB = NB;
    p->tiles = tiles;
    p->nrhs = nrhs;
    *rc = p->build();
    if (*rc != TTG_CHOL_OK) {
      plans.erase(key);
      return nullptr;
    }
  }
  *rc = TTG_CHOL_OK;
  return p.get();
}

template SolvePlan<double> *get_solve_plan<double>(int, int, const double *, int, int *);
template SolvePlan<float> *get_solve_plan<float>(int, int, const float *, int, int *);

// ---------------------------------------------------------------------------
// residual
// ---------------------------------------------------------------------------
constexpr int kRB = 64;  // LDS block of A

// X (column-major, ldx) columns j0 .. j0+nr-1 -> Xr (Np x kRhsChunk, row i at
// i * kRhsChunk; zero beyond N and beyond nr)
__global__ void residual_rows_kernel(int N, int Np, const double *__restrict__ X, long ldx, int nr,
                                     double *__restrict__ Xr) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)Np * kRhsChunk) return;
  const int j = (int)(g % kRhsChunk), i = (int)(g / kRhsChunk);
  Xr[g] = (i < N && j < nr) ? X[i + (long)j * ldx] : 0.0;
}

// One workgroup per lower tile (m, n):  P = A_mn X_n  (rows of block m) and
// Q = A_mn^T X_m (rows of block n; strictly lower part on a diagonal tile),
// both on FP64 MFMA from 64 x 64 blocks of A staged in LDS (each element of A
// read from HBM once), stored row-major to the tile's slots W[t][0|1][NB][16].
// Wave w owns rows w*16..w*16+15 of every 64-row block for P and of every
// 64-column block for Q.  The general form: any NB, diagonal and ragged
// tiles.  rest = 0: workgroup b takes tile b of the lower tiles; rest = 1:
// only the tiles residual_strip_kernel leaves (diagonal tiles b < nt, then
// the last block row's when N is not a multiple of NB).
__global__ __launch_bounds__(256) void residual_tile_kernel(int N, int NB, int nt, int rest,
                                                            const double *__restrict__ tiles,
                                                            const double *__restrict__ Xr, double *__restrict__ W) {
  int m, n;
  long t;
  if (rest) {
    const int b = blockIdx.x;
    m = b < nt ? b : nt - 1;
    n = b < nt ? b : b - nt;
    t = tile_index(m, n, nt);
  } else {
    t = blockIdx.x;
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
  const double *A = tiles + t * (long)NB * NB;
  double *W0 = W + t * 2L * kRhsChunk * NB;
  double *W1 = W0 + (long)kRhsChunk * NB;
  const double *Xn = Xr + (long)n * NB * kRhsChunk;
  const double *Xm = Xr + (long)m * NB * kRhsChunk;
  __shared__ double blk[kRB][kRB + 1];  // blk[c][r]
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  const bool diag = (m == n);
  const int nsb = (NB + kRB - 1) / kRB;
  const int mrows = min(NB, N - m * NB), ncols = min(NB, N - n * NB);
  d4_t pacc[8];  // nsb <= 8 (NB <= 512)
#pragma unroll
  for (int a = 0; a < 8; ++a) pacc[a] = d4_t{0.0, 0.0, 0.0, 0.0};
  // blocks (cs, rs) in column-major order; a diagonal tile has none above
  // its diagonal.  The next block's 16 elements per thread are loaded (from
  // clamped, always valid addresses) while this block is multiplied, and
  // masked when they are written to LDS.
  auto active = [&](int cs, int rs) { return rs < nsb && !(diag && rs < cs); };
  auto next_block = [&](int &cs, int &rs) {
    do {
      if (++rs == nsb) {
        rs = 0;
        ++cs;
      }
    } while (cs < nsb && !active(cs, rs));
  };
  double av[16];
  auto fetch = [&](int cs, int rs) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int e = tid + 256 * i, r = e & 63, c = e >> 6;
      const int gr = min(rs * kRB + r, mrows - 1), gc = min(cs * kRB + c, ncols - 1);
      av[i] = A[gr + (long)gc * NB];
    }
  };
  auto stash = [&](int cs, int rs) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int e = tid + 256 * i, r = e & 63, c = e >> 6;
      const int gr = rs * kRB + r, gc = cs * kRB + c;
      const bool ok = gr < mrows && gc < ncols && (!diag || gr >= gc);
      blk[c][r] = ok ? av[i] : 0.0;
    }
  };
  int ncs = 0, nrs = 0;  // the block whose elements are in av
  if (!active(0, 0)) next_block(ncs, nrs);
  fetch(ncs, nrs);
  for (int cs = 0; cs < nsb; ++cs) {
    // B operand of P: X_n rows cs*64 + 4 ks + lk, right-hand side lr
    double xb[kRB / 4];
#pragma unroll
    for (int ks = 0; ks < kRB / 4; ++ks) {
      const int k = cs * kRB + 4 * ks + lk;
      xb[ks] = (k < NB) ? Xn[(long)k * kRhsChunk + lr] : 0.0;
    }
    d4_t qacc = {0.0, 0.0, 0.0, 0.0};
    // a runtime loop over the row blocks (nothing of later blocks is hoisted);
    // the P accumulators rotate so that block rs is always pacc[0], and after
    // the 8 steps every block is back in its own slot
#pragma unroll 1
    for (int rs = 0; rs < 8; ++rs) {
      if (active(cs, rs)) {
        __syncthreads();
        stash(cs, rs);
        next_block(ncs, nrs);
        if (ncs < nsb) fetch(ncs, nrs);
        double xm[kRB / 4];
#pragma unroll
        for (int ks = 0; ks < kRB / 4; ++ks) {
          const int k = rs * kRB + 4 * ks + lk;
          xm[ks] = (k < NB) ? Xm[(long)k * kRhsChunk + lr] : 0.0;
        }
        __syncthreads();
        // P rows rs*64 + w*16 + (0..15):  sum over this block's 64 columns
#pragma unroll
        for (int ks = 0; ks < kRB / 4; ++ks)
          pacc[0] = Mfma<double>::mma(blk[4 * ks + lk][w * 16 + lr], xb[ks], pacc[0]);
        if (diag && rs == cs) {  // Q takes the strictly lower part only
          __syncthreads();
          if (tid < kRB) blk[tid][tid] = 0.0;
          __syncthreads();
        }
        // Q rows (= columns of A) cs*64 + w*16 + (0..15):  sum over this block's 64 rows
#pragma unroll
        for (int ks = 0; ks < kRB / 4; ++ks) qacc = Mfma<double>::mma(blk[w * 16 + lr][4 * ks + lk], xm[ks], qacc);
      }
      const d4_t first = pacc[0];
#pragma unroll
      for (int a = 0; a < 7; ++a) pacc[a] = pacc[a + 1];
      pacc[7] = first;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = cs * kRB + w * 16 + Mfma<double>::drow(lk, e);
      if (row < NB) W1[(long)row * kRhsChunk + lr] = qacc[e];
    }
  }
#pragma unroll
  for (int rs = 0; rs < 8; ++rs) {
    if (rs >= nsb) break;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = rs * kRB + w * 16 + Mfma<double>::drow(lk, e);
      if (row < NB) W0[(long)row * kRhsChunk + lr] = pacc[rs][e];
    }
  }
}

// The same products for the full off-diagonal tiles of NB = 512 (all but
// the diagonal tiles and a ragged last block row), written for HBM
// streaming.  A tile is read as 64 blocks of 16 columns x 256 rows (strip s
// = block / 2, half h = block % 2): each column of a block is 2 KB
// contiguous in HBM.  The blocks go HBM -> LDS by LDS-DMA loads through a
// ring of 4 stages (3 blocks in flight), one barrier per block.  Each of
// the 8 waves takes 32 rows of a block for both products: P (rows 256 h +
// 32 w + 16 i + (0..15), i = 0, 1; 4 k-slices of the strip's columns) and
// Q (the strip's 16 columns; 8 k-slices of its 32 rows, whose X_m operands
// stay in registers for the whole tile).  The operands of the next block
// are read into registers while this block's MFMAs run.  P accumulates
// over the columns in the order of residual_tile_kernel (bitwise the same
// P); Q of a strip is the sum of the 8 waves' partials in wave order.
constexpr int kStripNB = 512;
constexpr int kStripLd = 257;                          // LDS column stride of a block (doubles)
constexpr int kStripXn = 16 * kStripLd;                // X_n rows of the strip (16 x 16)
constexpr int kStripStage = kStripXn + 16 * kRhsChunk;  // doubles per stage
constexpr int kStripLoads = 17;                        // LDS-DMA loads per wave per block
constexpr int kStripBlocks = 2 * (kStripNB / 16);      // 64 blocks per tile

struct StripOps {
  double a[20];  // P operands of A [i][ks] (8), Q operands of A [ks] (8), X_n operands [ks] (4)
};

__global__ __launch_bounds__(512) void residual_strip_kernel(int nt, int nfull, const double *__restrict__ tiles,
                                                             const double *__restrict__ Xr, double *__restrict__ W) {
  // workgroup b -> full off-diagonal tile (m, n), n < m < nfull, column by column
  const int b = blockIdx.x;
  int n, m;
  {
    int lo = 0, hi = nfull - 2;
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (mid * (nfull - 1) - mid * (mid - 1) / 2 <= b)
        lo = mid;
      else
        hi = mid - 1;
    }
    n = lo;
    m = n + 1 + (b - (n * (nfull - 1) - n * (n - 1) / 2));
  }
  const long t = tile_index(m, n, nt);
  const double *A = tiles + t * (long)kStripNB * kStripNB;
  double *W0 = W + t * 2L * kRhsChunk * kStripNB;
  double *W1 = W0 + (long)kRhsChunk * kStripNB;
  const double *Xn = Xr + (long)n * kStripNB * kRhsChunk;
  const double *Xm = Xr + (long)m * kStripNB * kRhsChunk;
  __shared__ __align__(16) double lds[4][kStripStage];
  __shared__ __align__(16) double qpart[8][256];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  // block j into stage st: wave w loads rows 32 w .. 32 w + 31 of the 16
  // columns (one 4-byte word per lane, 256 bytes per load) and rows 16 s +
  // 2 w, + 1 of X_n; the bases are scalar, lane l loads bytes 4 l .. 4 l + 3
  auto issue = [&](int j, int st) {
    const int s = j >> 1, h = j & 1;
    double *sb = lds[st];
    const char *src = reinterpret_cast<const char *>(A + (long)(16 * s) * kStripNB + 256 * h + 32 * w);
    const unsigned voff = 4u * (unsigned)lane;
#pragma unroll
    for (int c = 0; c < 16; ++c) lds_dma_dword_s(src + (long)c * kStripNB * 8, voff, sb + c * kStripLd + 32 * w);
    lds_dma_dword_s(Xn + (long)(16 * s + 2 * w) * kRhsChunk, voff, sb + kStripXn + 32 * w);
  };
  auto read = [&](int st, StripOps &o) {
    const double *sb = lds[st];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) o.a[4 * i + ks] = sb[(4 * ks + lk) * kStripLd + 32 * w + 16 * i + lr];
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) o.a[8 + ks] = sb[lr * kStripLd + 32 * w + 4 * ks + lk];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) o.a[16 + ks] = sb[kStripXn + (4 * ks + lk) * kRhsChunk + lr];
  };
  d4_t pacc[2][2];  // [h][i]
  d4_t qacc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int i = 0; i < 2; ++i) pacc[h][i] = d4_t{0.0, 0.0, 0.0, 0.0};
  double xq[2][8];  // X_m rows 256 h + 32 w + 4 ks + lk
  auto mma = [&](int h, const StripOps &o) {
    auto body = [&](d4_t &c0, d4_t &c1, const double(&x)[8]) {
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        c0 = Mfma<double>::mma(o.a[ks], o.a[16 + ks], c0);
        c1 = Mfma<double>::mma(o.a[4 + ks], o.a[16 + ks], c1);
        qacc = Mfma<double>::mma(o.a[8 + 2 * ks], x[2 * ks], qacc);
        qacc = Mfma<double>::mma(o.a[9 + 2 * ks], x[2 * ks + 1], qacc);
      }
    };
    if (h == 0)
      body(pacc[0][0], pacc[0][1], xq[0]);
    else
      body(pacc[1][0], pacc[1][1], xq[1]);
  };
  // Q of strip s: the 8 partials summed in wave order, 32 elements per wave
  auto q_sum = [&](int s) {
    if (lane < 32) {
      const int e = 32 * w + lane, pl = e & 63, pr = e >> 6;
      double v = 0.0;
#pragma unroll
      for (int q = 0; q < 8; ++q) v += qpart[q][pl * 4 + pr];
      W1[(long)(16 * s + Mfma<double>::drow(pl >> 4, pr)) * kRhsChunk + (pl & 15)] = v;
    }
  };
  // blocks 0 .. 3 in flight, then this wave's X_m operands straight into
  // registers; one wait for all
  for (int j = 0; j < 4; ++j) issue(j, j);
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) xq[h][ks] = Xm[(long)(256 * h + 32 * w + 4 * ks + lk) * kRhsChunk + lr];
  wait_vmcnt<0>();
  raw_barrier();
  StripOps ops[2];
  read(0, ops[0]);
  // Block j (operands in cur): wait for this wave's reads of block j's stage
  // and for its loads of block j + 1 (blocks j + 2 and j + 3 may be
  // outstanding; the W1 stores of q_sum go out before each iteration's
  // loads, so at least those 34 loads are younger than block j + 1's), then
  // one barrier publishes block j + 1 and frees block j's stage for j + 4.
  auto step = [&](int j, StripOps &cur, StripOps &nxt) {
    wait_lgkmcnt0();
    if (j + 3 < kStripBlocks)
      wait_vmcnt<2 * kStripLoads>();
    else if (j + 2 < kStripBlocks)
      wait_vmcnt<kStripLoads>();
    else
      wait_vmcnt<0>();
    raw_barrier();
    const int h = j & 1;
    if (j >= 2 && h == 0) q_sum((j >> 1) - 1);
    if (j + 4 < kStripBlocks) issue(j + 4, j & 3);
    if (j + 1 < kStripBlocks) read((j + 1) & 3, nxt);
    mma(h, cur);
    if (h == 1) {  // the strip ends: this wave's Q partial
#pragma unroll
      for (int r = 0; r < 4; ++r) qpart[w][lane * 4 + r] = qacc[r];
      qacc = d4_t{0.0, 0.0, 0.0, 0.0};
    }
  };
  for (int j = 0; j < kStripBlocks; j += 2) {
    step(j, ops[0], ops[1]);
    step(j + 1, ops[1], ops[0]);
  }
  wait_lgkmcnt0();
  raw_barrier();
  q_sum(kStripBlocks / 2 - 1);
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        W0[(long)(256 * h + 32 * w + 16 * i + Mfma<double>::drow(lk, e)) * kRhsChunk + lr] = pacc[h][i][e];
}

// R(i, j) = B(i, j) - sum_{n <= m} P(m, n)[r][j] - sum_{n' >= m} Q(n', m)[r][j],
// m = i / NB, r = i % NB: a fixed summation order (deterministic)
__global__ void residual_reduce_kernel(int N, int NB, int nt, const double *__restrict__ W, int nr,
                                       const double *__restrict__ B, long ldb, double *__restrict__ R, long ldr,
                                       int j0) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)N * nr) return;
  const int j = (int)(g % nr), i = (int)(g / nr);
  const int m = i / NB, r = i % NB;
  double s = 0.0;
  for (int n = 0; n <= m; ++n) {
    const long t = tile_index(m, n, nt);
    s += W[t * 2L * kRhsChunk * NB + (long)r * kRhsChunk + j];
  }
  for (int n2 = m; n2 < nt; ++n2) {
    const long t = tile_index(n2, m, nt);
    s += W[t * 2L * kRhsChunk * NB + (long)kRhsChunk * NB + (long)r * kRhsChunk + j];
  }
  R[i + (long)(j0 + j) * ldr] = B[i + (long)(j0 + j) * ldb] - s;
}

static std::mutex g_res_mu;
static std::map<hipStream_t, std::unique_ptr<DevBuf>> g_res_ws;

void release_solve_plans() {
  {
    std::lock_guard<std::mutex> lk(g_solve_mu);
    g_dplans.clear();
    g_splans.clear();
  }
  std::lock_guard<std::mutex> lk(g_res_mu);
  for (auto &kv : g_res_ws)
    if (kv.second && kv.second->ptr) (void)hipFree(kv.second->ptr);
  g_res_ws.clear();
}

int residual_fp64(hipStream_t s, int N, int NB, const double *tiles, int nrhs, const double *X, long ldx,
                  const double *B, long ldb, double *R, long ldr) {
  const int nt = (N + NB - 1) / NB;
  const long Np = (long)nt * NB;
  const long ntiles = (long)nt * (nt + 1) / 2;
  DevBuf *ws;
  {
    std::lock_guard<std::mutex> lk(g_res_mu);
    auto &p = g_res_ws[s];
    if (!p) p.reset(new DevBuf);
    ws = p.get();
  }
  const size_t wsz = (size_t)ntiles * 2 * kRhsChunk * NB;
  // block rows of full NB = 512 tiles (the strip kernel's), 0 otherwise
  const int nfull = NB == kStripNB ? N / NB : 0;
  TTGC_HIP(ws->reserve(sizeof(double) * (wsz + (size_t)Np * kRhsChunk)));
  double *W = static_cast<double *>(ws->ptr);
  double *Xr = W + wsz;
  for (int j0 = 0; j0 < nrhs; j0 += kRhsChunk) {
    const int nr = std::min(kRhsChunk, nrhs - j0);
    hipLaunchKernelGGL(residual_rows_kernel, dim3((unsigned)((Np * kRhsChunk + 255) / 256)), dim3(256), 0, s, N,
                       (int)Np, X + (long)j0 * ldx, ldx, nr, Xr);
    TTGC_HIP(hipGetLastError());
    if (nfull >= 2) {
      // full off-diagonal tiles on the strip kernel, the others on the general one
      hipLaunchKernelGGL(residual_strip_kernel, dim3((unsigned)(nfull * (nfull - 1) / 2)), dim3(512), 0, s, nt, nfull,
                         tiles, Xr, W);
      TTGC_HIP(hipGetLastError());
      hipLaunchKernelGGL(residual_tile_kernel, dim3((unsigned)(nt + (nfull < nt ? nt - 1 : 0))), dim3(256), 0, s, N,
                         NB, nt, 1, tiles, Xr, W);
    } else {
      hipLaunchKernelGGL(residual_tile_kernel, dim3((unsigned)ntiles), dim3(256), 0, s, N, NB, nt, 0, tiles, Xr, W);
    }
    TTGC_HIP(hipGetLastError());
    const long n = (long)N * nr;
    hipLaunchKernelGGL(residual_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, N, NB, nt, W, nr,
                       B, ldb, R, ldr, j0);
    TTGC_HIP(hipGetLastError());
  }
  return TTG_CHOL_OK;
}

}  // namespace ttgc

using namespace ttgc;

extern "C" {

int ttg_chol_dpotrs_tiled(void *stream, int N, int NB, const double *tiles, int nrhs, double *B, int ldb) {
  clear_error();
  if (N < 0 || NB <= 0 || NB > TTG_CHOL_MAX_TILE || NB % 32 != 0 || nrhs < 0 || ldb < std::max(1, N) ||
      (N > 0 && nrhs > 0 && (!tiles || !B))) {
    set_error("dpotrs_tiled: bad argument (N=%d NB=%d nrhs=%d ldb=%d)", N, NB, nrhs, ldb);
    return TTG_CHOL_EINVAL;
  }
  if (N == 0 || nrhs == 0) return TTG_CHOL_OK;
  if (nrhs > kSweepMaxRhs) {
    // the sweeps take up to 16 right-hand sides per pass: split wide solves
    for (int j0 = 0; j0 < nrhs; j0 += kSweepMaxRhs) {
      int rc = ttg_chol_dpotrs_tiled(stream, N, NB, tiles, std::min(kSweepMaxRhs, nrhs - j0), B + (long)j0 * ldb,
                                     ldb);
      if (rc != TTG_CHOL_OK) return rc;
    }
    return TTG_CHOL_OK;
  }
  hipStream_t s = as_stream(stream);
  int rc = 0;
  SolvePlan<double> *p = get_solve_plan<double>(N, NB, tiles, nrhs, &rc);
  if (!p) return rc;
  double *bt = static_cast<double *>(p->d_bt.ptr);
  rc = p->prepare(s);
  if (rc != TTG_CHOL_OK) return rc;
  TTGC_HIP(launch_to_rows<double, double>(s, N, p->Np, nrhs, B, ldb, bt));
  rc = p->run_sweeps(s);
  if (rc != TTG_CHOL_OK) return rc;
  TTGC_HIP(launch_from_rows<double, double>(s, N, nrhs, bt, B, ldb));
  return TTG_CHOL_OK;
}

int ttg_chol_spotrs_tiled(void *stream, int N, int NB, const float *tiles, int nrhs, double *B, int ldb) {
  clear_error();
  if (N < 0 || NB <= 0 || NB > TTG_CHOL_MAX_TILE || NB % 32 != 0 || nrhs < 0 || ldb < std::max(1, N) ||
      (N > 0 && nrhs > 0 && (!tiles || !B))) {
    set_error("spotrs_tiled: bad argument (N=%d NB=%d nrhs=%d ldb=%d)", N, NB, nrhs, ldb);
    return TTG_CHOL_EINVAL;
  }
  if (N == 0 || nrhs == 0) return TTG_CHOL_OK;
  if (nrhs > kSweepMaxRhs) {
    for (int j0 = 0; j0 < nrhs; j0 += kSweepMaxRhs) {
      int rc = ttg_chol_spotrs_tiled(stream, N, NB, tiles, std::min(kSweepMaxRhs, nrhs - j0), B + (long)j0 * ldb,
                                     ldb);
      if (rc != TTG_CHOL_OK) return rc;
    }
    return TTG_CHOL_OK;
  }
  hipStream_t s = as_stream(stream);
  int rc = 0;
  SolvePlan<float> *p = get_solve_plan<float>(N, NB, tiles, nrhs, &rc);
  if (!p) return rc;
  rc = p->prepare(s);
  if (rc != TTG_CHOL_OK) return rc;
  TTGC_HIP(launch_to_rows<double, float>(s, N, p->Np, nrhs, B, ldb, p->bt()));
  rc = p->run_sweeps(s);
  if (rc != TTG_CHOL_OK) return rc;
  TTGC_HIP(launch_from_rows<float, double>(s, N, nrhs, p->bt(), B, ldb));
  return TTG_CHOL_OK;
}

int ttg_chol_dresidual_tiled(void *stream, int N, int NB, const double *tiles, int nrhs, const double *X, int ldx,
                             const double *B, int ldb, double *R, int ldr) {
  clear_error();
  if (N < 0 || NB <= 0 || NB > TTG_CHOL_MAX_TILE || nrhs < 0 || ldx < std::max(1, N) || ldb < std::max(1, N) ||
      ldr < std::max(1, N) || (N > 0 && nrhs > 0 && (!tiles || !X || !B || !R))) {
    set_error("dresidual_tiled: bad argument (N=%d NB=%d nrhs=%d)", N, NB, nrhs);
    return TTG_CHOL_EINVAL;
  }
  if (N == 0 || nrhs == 0) return TTG_CHOL_OK;
  return residual_fp64(as_stream(stream), N, NB, tiles, nrhs, X, ldx, B, ldb, R, ldr);
}

}  // extern "C"
