// Mixed-precision solve with FP64 iterative refinement (no reference
// counterpart; the target of the tiled factorization on MI355X).
//
//   1. A (FP64 lower tiles, kept intact) is copied into a working factor in
//      the working precision (FP32, or FP64 for the plain path).
//   2. The tiled factorization of the examples/potrf graph runs on it: the
//      diagonal POTRF always computes in FP64 (tile converted on the fly),
//      TRSM and the SYRK/GEMM trailing updates in the working precision.
//   3. X0 = (L L^T)^{-1} B in the working precision.
//   4. Refinement in FP64:  R = B - A X (FP64, on the original A), stop when
//      max_j ||r_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf) <= tol,
//      else D = (L L^T)^{-1} R (working precision) and X += D (FP64).
// Or (ttg_chol_dposv_krylov) preconditioned conjugate gradients in FP64 on
// A x = b with M = L L^T, the low-precision factor of A (optionally of A
// with its diagonal scaled by 1 + shift, for factors that would otherwise
// meet a non-positive pivot), one CG per right-hand side.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#include "chol_f16.hpp"
#include "chol_internal.hpp"
#include "chol_kernels.hpp"
#include "chol_solve.hpp"
#include "chol_tiled.hpp"
#include "chol_tiled_plan.hpp"

namespace ttgc {

// Working copy and ||A||_inf in one pass over the FP64 tiles: one workgroup
// per lower tile (m, n) converts it and writes |A| partial sums, S[t][0][r] =
// sum over the tile's row r (lower part on a diagonal tile) and S[t][1][c] =
// sum over its column c (strictly lower part on a diagonal tile: the
// transpose's contribution to row n*NB + c).  Thread t owns rows t and
// t + 256 (NB <= 512); each wave reduces its rows of a column by shuffles in
// a fixed order, and the four wave sums are added in wave order.
template <typename T>
__global__ __launch_bounds__(256) void convert_norm_kernel(int N, int NB, int nt, long t0, const double *__restrict__ A,
                                                           T *__restrict__ W, double *__restrict__ S,
                                                           double dscale) {
  const long t = t0 + blockIdx.x;
  int m, n;
  {
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
  const double *a = A + t * (long)NB * NB;
  T *w = W + t * (long)NB * NB;
  double *rs = S + t * 2L * NB, *cs = rs + NB;
  __shared__ double csum[4][512];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int mrows = min(NB, N - m * NB), ncols = min(NB, N - n * NB);
  const bool diag = m == n;
  const int r0 = tid, r1 = tid + 256;
  double ra0 = 0.0, ra1 = 0.0;
  for (int c = 0; c < NB; ++c) {
    const long o = (long)c * NB;
    double u = 0.0;
    if (r0 < NB) {
      const double v = a[o + r0];
      w[o + r0] = (T)(diag && r0 == c ? v * dscale : v);
      const bool in = r0 < mrows && c < ncols;
      if (in && (!diag || r0 >= c)) ra0 += fabs(v);
      if (in && (!diag || r0 > c)) u += fabs(v);
    }
    if (r1 < NB) {
      const double v = a[o + r1];
      w[o + r1] = (T)(diag && r1 == c ? v * dscale : v);
      const bool in = r1 < mrows && c < ncols;
      if (in && (!diag || r1 >= c)) ra1 += fabs(v);
      if (in && (!diag || r1 > c)) u += fabs(v);
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

// -rowsum_i of |A| from the tile partial sums, in a fixed order (row block m:
// tiles (m, 0..m) by rows, then tiles (m..nt-1, m) by columns).
__global__ void norm_rows_kernel(int N, int NB, int nt, const double *__restrict__ S, double *__restrict__ rows) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int m = i / NB, r = i % NB;
  double s = 0.0;
  for (int n = 0; n <= m; ++n) s += S[tile_index(m, n, nt) * 2L * NB + r];
  for (int n2 = m; n2 < nt; ++n2) s += S[tile_index(n2, m, nt) * 2L * NB + NB + r];
  rows[i] = -s;
}

// Per-column reductions over N rows, spread over kRedParts workgroups per
// column (rows in kRedParts contiguous chunks) and combined per column in
// chunk order by a second launch: fixed summation order, and the whole GPU
// on a pass that one workgroup per column left latency-bound.
constexpr int kRedParts = 32;

// max |R_j|, max |X_j|, max |B_j| of chunk g of column j (NaN in R propagates)
__global__ __launch_bounds__(256) void berr_part_kernel(int N, const double *__restrict__ R, long ldr,
                                                        const double *__restrict__ X, long ldx,
                                                        const double *__restrict__ B, long ldb,
                                                        double *__restrict__ part) {
  const int g = blockIdx.x, j = blockIdx.y;
  const int chunk = (N + kRedParts - 1) / kRedParts, i0 = g * chunk, i1 = min(N, i0 + chunk);
  __shared__ double sr[256], sx[256], sb[256];
  double r = 0.0, x = 0.0, b = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += 256) {
    const double rv = fabs(R[i + (long)j * ldr]);
    r = (rv > r || rv != rv) ? rv : r;
    x = fmax(x, fabs(X[i + (long)j * ldx]));
    b = fmax(b, fabs(B[i + (long)j * ldb]));
  }
  sr[threadIdx.x] = r;
  sx[threadIdx.x] = x;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double o = sr[threadIdx.x + w];
      sr[threadIdx.x] = (o > sr[threadIdx.x] || o != o) ? o : sr[threadIdx.x];
      sx[threadIdx.x] = fmax(sx[threadIdx.x], sx[threadIdx.x + w]);
      sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double *o = part + 3L * ((long)j * kRedParts + g);
    o[0] = sr[0];
    o[1] = sx[0];
    o[2] = sb[0];
  }
}

// Per right-hand side j: berr_j = max|R_j| / (anorm * max|X_j| + max|B_j|).
__global__ void berr_final_kernel(int nrhs, const double *__restrict__ part, const double *__restrict__ anorm,
                                  double *__restrict__ out) {
  for (int j = threadIdx.x; j < nrhs; j += blockDim.x) {
    double r = 0.0, x = 0.0, b = 0.0;
    for (int g = 0; g < kRedParts; ++g) {
      const double *o = part + 3L * ((long)j * kRedParts + g);
      r = (o[0] > r || o[0] != o[0]) ? o[0] : r;
      x = fmax(x, o[1]);
      b = fmax(b, o[2]);
    }
    out[j] = r / (anorm[0] * x + b);
  }
}

// anorm = max_i (-rowsum_i) where rowsum = 0 - |A| 1
__global__ __launch_bounds__(1024) void max_neg_kernel(int N, const double *__restrict__ v, double *__restrict__ out) {
  __shared__ double s[1024];
  double m = 0.0;
  for (int i = threadIdx.x; i < N; i += 1024) m = fmax(m, -v[i]);
  s[threadIdx.x] = m;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

// X(i, j) (+)= Bt(j, i) for the first N rows (Bt holds nrhs x Np, transposed).
template <typename TS, bool ADD>
__global__ void from_rows_acc_kernel(int N, int nrhs, const TS *__restrict__ Bt, double *__restrict__ X, long ldx) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)N * nrhs) return;
  const int i = (int)(g % N), j = (int)(g / N);
  const double v = (double)Bt[j + (long)i * nrhs];
  if (ADD)
    X[i + (long)j * ldx] += v;
  else
    X[i + (long)j * ldx] = v;
}

// sum over chunk g of column j of X(i, j) Y(i, j), reduced in a fixed order
__global__ __launch_bounds__(256) void dot_part_kernel(int N, const double *__restrict__ X, long ldx,
                                                       const double *__restrict__ Y, long ldy,
                                                       double *__restrict__ part) {
  const int g = blockIdx.x, j = blockIdx.y;
  const int chunk = (N + kRedParts - 1) / kRedParts, i0 = g * chunk, i1 = min(N, i0 + chunk);
  __shared__ double sm[256];
  double a = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += 256) a += X[i + (long)j * ldx] * Y[i + (long)j * ldy];
  sm[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long)j * kRedParts + g] = sm[0];
}

// out[j] = sum of column j's chunk sums, in chunk order
__global__ void dot_final_kernel(int nrhs, const double *__restrict__ part, double *__restrict__ out) {
  for (int j = threadIdx.x; j < nrhs; j += blockDim.x) {
    double a = 0.0;
    for (int g = 0; g < kRedParts; ++g) a += part[(long)j * kRedParts + g];
    out[j] = a;
  }
}

// CG step of every column j, with Q = -A P:  a_j = rz_j / (p_j . A p_j),
// X += a P, R -= a A P  (a = 0 where p_j . A p_j = 0)
__global__ void cg_step_kernel(int N, int nrhs, const double *__restrict__ rz, const double *__restrict__ pq,
                               const double *__restrict__ P, double *__restrict__ X, long ldx,
                               const double *__restrict__ Q, double *__restrict__ R) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)N * nrhs) return;
  const int i = (int)(g % N), j = (int)(g / N);
  const double pap = -pq[j];
  const double a = pap != 0.0 ? rz[j] / pap : 0.0;
  X[i + (long)j * ldx] += a * P[g];
  R[g] += a * Q[g];
}

// P := Z + (rz_new_j / rz_j) P  (beta = 0 where rz_j = 0)
__global__ void cg_dir_kernel(int N, int nrhs, const double *__restrict__ rz_new, const double *__restrict__ rz,
                              const double *__restrict__ Z, double *__restrict__ P) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)N * nrhs) return;
  const int j = (int)(g / N);
  const double b = rz[j] != 0.0 ? rz_new[j] / rz[j] : 0.0;
  P[g] = Z[g] + b * P[g];
}

struct MixedPlanBase {
  virtual ~MixedPlanBase() = default;
};

template <typename T>
struct MixedPlan : MixedPlanBase {
  int N = 0, NB = 0, nrhs = 0;
  int lowp = 0;  // 0: FP32 updates, else kLowF16x2 / kLowF16 / kLowBF16
  DevBuf W;      // working factor tiles (T)
  DevBuf R;      // residual, N x nrhs (FP64)
  DevBuf stats;  // anorm, berr[nrhs], then the per-chunk partials of the column reductions
  DevBuf rows, norm_parts;  // -rowsum of |A|; per-tile |A| partial sums
  DevBuf cg, zeros;         // CG: P, Q, Z (N x nrhs each) and per-column scalars; a zero N x nrhs
  std::unique_ptr<TiledPlanT<T>> fplan;
  SolvePlan<T> *splan = nullptr;
  std::vector<double> hstats;

  ~MixedPlan() override {
    for (DevBuf *b : {&W, &R, &stats, &rows, &norm_parts, &cg, &zeros})
      if (b->ptr) (void)hipFree(b->ptr);
  }

  T *w() { return static_cast<T *>(W.ptr); }
  double *r() { return static_cast<double *>(R.ptr); }
  double *st() { return static_cast<double *>(stats.ptr); }
  double *red_part() { return st() + 1 + nrhs; }

  int build() {
    const int nt = (N + NB - 1) / NB;
    const size_t ntile = (size_t)nt * (nt + 1) / 2 * NB * NB;
    TTGC_HIP(W.reserve(sizeof(T) * ntile));
    TTGC_HIP(R.reserve(sizeof(double) * (size_t)N * nrhs));
    TTGC_HIP(stats.reserve(sizeof(double) * (1 + nrhs + 3 * (size_t)kRedParts * nrhs)));
    TTGC_HIP(rows.reserve(sizeof(double) * (size_t)N));
    TTGC_HIP(norm_parts.reserve(sizeof(double) * (size_t)nt * (nt + 1) / 2 * 2 * NB));
    fplan.reset(new TiledPlanT<T>);
    fplan->N = N;
    fplan->NB = NB;
    fplan->tiles = w();
    fplan->flags = TTG_CHOL_F_DEFAULT;
    fplan->lowp = lowp;
    int rc = fplan->build();
    if (rc) return rc;
    splan = get_solve_plan<T>(N, NB, w(), nrhs, &rc);
    if (!splan) return rc;
    hstats.resize(1 + nrhs);
    return TTG_CHOL_OK;
  }

  // Solve into X (overwrite when !add, accumulate otherwise) the rhs Y.
  int solve(hipStream_t s, const double *Y, long ldy, double *X, long ldx, bool add) {
    T *bt = splan->bt();
    TTGC_HIP(launch_to_rows<double, T>(s, N, splan->Np, nrhs, Y, ldy, bt));
    int rc = splan->run_sweeps(s);
    if (rc) return rc;
    const long n = (long)N * nrhs;
    if (add)
      hipLaunchKernelGGL((from_rows_acc_kernel<T, true>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, N, nrhs,
                         bt, X, ldx);
    else
      hipLaunchKernelGGL((from_rows_acc_kernel<T, false>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, N,
                         nrhs, bt, X, ldx);
    TTGC_HIP(hipGetLastError());
    return TTG_CHOL_OK;
  }

  // The working factor as FP64 lower tiles (the tile layout, padding as
  // the factorization prepared it).
  int factor(hipStream_t s, double *L) const {
    const int nt = (N + NB - 1) / NB;
    const long n = (long)nt * (nt + 1) / 2 * NB * NB;
    TTGC_HIP(launch_widen<T>(s, n, static_cast<const T *>(W.ptr), L));
    return TTG_CHOL_OK;
  }

  // Synchronise s; an error if the last persistent sweep gave up (the
  // solves run before every synchronisation of the refinement loops).
  int sync_checked(hipStream_t s) {
    int failed = 0;
    if (const int *fw = splan->fail_word())
      TTGC_HIP(hipMemcpyAsync(&failed, fw, sizeof(int), hipMemcpyDeviceToHost, s));
    TTGC_HIP(hipStreamSynchronize(s));
    if (failed) {
      set_error("triangular sweep: a workgroup waited past the time limit");
      return TTG_CHOL_EHIP;
    }
    return TTG_CHOL_OK;
  }

  // st()[1 + j] = backward error of column j (N rows; R, X and B column-major)
  int berr_launch(hipStream_t s, const double *R, long ldr, const double *X, long ldx, const double *B, long ldb) {
    double *part = red_part();
    hipLaunchKernelGGL(berr_part_kernel, dim3(kRedParts, nrhs), dim3(256), 0, s, N, R, ldr, X, ldx, B, ldb, part);
    TTGC_HIP(hipGetLastError());
    hipLaunchKernelGGL(berr_final_kernel, dim3(1), dim3(256), 0, s, nrhs, part, st(), st() + 1);
    TTGC_HIP(hipGetLastError());
    return TTG_CHOL_OK;
  }
  // out[j] = sum_i X(i, j) Y(i, j) over the N rows (fixed order)
  int dot_cols(hipStream_t s, const double *X, long ldx, const double *Y, long ldy, double *out) {
    double *part = red_part();
    hipLaunchKernelGGL(dot_part_kernel, dim3(kRedParts, nrhs), dim3(256), 0, s, N, X, ldx, Y, ldy, part);
    TTGC_HIP(hipGetLastError());
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(256), 0, s, nrhs, part, out);
    TTGC_HIP(hipGetLastError());
    return TTG_CHOL_OK;
  }

  // backward error of every column from the residual in `res` (FP64), the
  // largest over the columns into *e (NaN propagates); host-synchronising
  int column_berr(hipStream_t s, const double *res, const double *B, long ldb, const double *X, long ldx,
                  double *e) {
    int rc = berr_launch(s, res, (long)N, X, ldx, B, ldb);
    if (rc) return rc;
    TTGC_HIP(hipMemcpyAsync(hstats.data(), st(), sizeof(double) * (1 + nrhs), hipMemcpyDeviceToHost, s));
    rc = sync_checked(s);
    if (rc) return rc;
    double m = 0.0;
    for (int j = 0; j < nrhs; ++j) m = (hstats[1 + j] > m || std::isnan(hstats[1 + j])) ? hstats[1 + j] : m;
    *e = m;
    return TTG_CHOL_OK;
  }

  // Working copy of A (diagonal scaled by dscale) with the tiles' row and
  // column sums of |A|, -||A||_inf into st()[0], and the factorization of
  // the copy.  The tiles of the columns the first panel chain works on are
  // converted first, on s; the others, and the norm, go to the plan's bulk
  // stream, where they run beside that chain.
  int convert_and_factor(hipStream_t s, const double *A, double dscale, int *d_info) {
    const int nt = (N + NB - 1) / NB;
    const long ntiles = (long)nt * (nt + 1) / 2;
    const int c = fplan->first_cols();
    const long t1 = (long)c * nt - (long)c * (c - 1) / 2;  // tiles of columns < c
    double *parts = static_cast<double *>(norm_parts.ptr), *rw = static_cast<double *>(rows.ptr);
    T *wk = w();
    double *stat = st();
    const int n = N, nb = NB;
    auto convert = [=](hipStream_t q, long a, long b) -> int {
      if (b <= a) return TTG_CHOL_OK;
      hipLaunchKernelGGL(convert_norm_kernel<T>, dim3((unsigned)(b - a)), dim3(256), 0, q, n, nb, nt, a, A, wk,
                         parts, dscale);
      TTGC_HIP(hipGetLastError());
      return TTG_CHOL_OK;
    };
    auto norm = [=](hipStream_t q) -> int {
      hipLaunchKernelGGL(norm_rows_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, q, n, nb, nt, parts, rw);
      TTGC_HIP(hipGetLastError());
      hipLaunchKernelGGL(max_neg_kernel, dim3(1), dim3(1024), 0, q, n, rw, stat);
      TTGC_HIP(hipGetLastError());
      return TTG_CHOL_OK;
    };
    int rc = convert(s, 0, t1);
    if (rc) return rc;
    if (t1 < ntiles) {
      fplan->pre_bulk = [=](hipStream_t q) -> int {
        const int r = convert(q, t1, ntiles);
        return r ? r : norm(q);
      };
    } else {
      rc = norm(s);
      if (rc) return rc;
    }
    rc = fplan->run(s, d_info);
    fplan->pre_bulk = nullptr;
    return rc;
  }

  // steps 1-3: working copy (diagonal scaled by dscale), ||A||_inf, factor,
  // X = M^{-1} B.  *failed: the factorization met a non-positive pivot.
  int factor_and_solve(hipStream_t s, const double *A, const double *B, long ldb, double *X, long ldx, int *d_info,
                       double dscale, bool *failed) {
    *failed = false;
    int rc = convert_and_factor(s, A, dscale, d_info);
    if (rc) return rc;
    int hinfo = 0;
    TTGC_HIP(hipMemcpyAsync(&hinfo, d_info, sizeof(int), hipMemcpyDeviceToHost, s));
    TTGC_HIP(hipStreamSynchronize(s));
    if (hinfo != 0) {
      *failed = true;
      return TTG_CHOL_OK;
    }
    rc = splan->prepare(s);
    if (rc) return rc;
    return solve(s, B, ldb, X, ldx, false);
  }

  // Preconditioned CG from X = M^{-1} B: iterations counted in *iters (one
  // matrix product and one preconditioner solve each).  Convergence is
  // judged on the recursively updated residual and confirmed on the true
  // one (R = B - A X); a confirmation that fails restarts CG from the true
  // residual.
  int run_pcg(hipStream_t s, const double *A, const double *B, long ldb, double *X, long ldx, double tol,
              int max_iter, int *d_info, double shift, int *iters, double *berr) {
    const long n = (long)N * nrhs;
    TTGC_HIP(cg.reserve(sizeof(double) * (3 * n + 3 * nrhs)));
    TTGC_HIP(zeros.reserve(sizeof(double) * n));
    double *P = static_cast<double *>(cg.ptr), *Q = P + n, *Z = Q + n, *rz = Z + n, *rzn = rz + nrhs,
           *pq = rzn + nrhs;
    double *zero = static_cast<double *>(zeros.ptr);
    TTGC_HIP(hipMemsetAsync(zero, 0, sizeof(double) * n, s));
    bool failed = false;
    int rc = factor_and_solve(s, A, B, ldb, X, ldx, d_info, 1.0 + shift, &failed);
    if (rc) return rc;
    *iters = 0;
    if (failed) {
      *berr = INFINITY;
      return TTG_CHOL_OK;  // caller reads *d_info
    }
    const unsigned gv = (unsigned)((n + 255) / 256);
    double e = INFINITY;
    int it = 0;
    bool restart = true;
    for (;;) {
      if (restart) {  // true residual, then a fresh search direction
        rc = residual_fp64(s, N, NB, A, nrhs, X, ldx, B, ldb, r(), N);
        if (rc) return rc;
        rc = column_berr(s, r(), B, ldb, X, ldx, &e);
        if (rc) return rc;
        if (!(e > tol) || it >= max_iter) break;
        rc = solve(s, r(), N, Z, N, false);
        if (rc) return rc;
        rc = dot_cols(s, r(), (long)N, Z, (long)N, rz);
        if (rc) return rc;
        TTGC_HIP(hipMemcpyAsync(P, Z, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        restart = false;
      }
      // Q = -A P
      rc = residual_fp64(s, N, NB, A, nrhs, P, N, zero, N, Q, N);
      if (rc) return rc;
      rc = dot_cols(s, P, (long)N, Q, (long)N, pq);
      if (rc) return rc;
      hipLaunchKernelGGL(cg_step_kernel, dim3(gv), dim3(256), 0, s, N, nrhs, rz, pq, P, X, ldx, Q, r());
      TTGC_HIP(hipGetLastError());
      ++it;
      double er = INFINITY;
      rc = column_berr(s, r(), B, ldb, X, ldx, &er);
      if (rc) return rc;
      if (!(er > tol) || std::isnan(er) || it >= max_iter) {
        restart = true;  // confirm on the true residual (or stop)
        continue;
      }
      rc = solve(s, r(), N, Z, N, false);
      if (rc) return rc;
      rc = dot_cols(s, r(), (long)N, Z, (long)N, rzn);
      if (rc) return rc;
      hipLaunchKernelGGL(cg_dir_kernel, dim3(gv), dim3(256), 0, s, N, nrhs, rzn, rz, Z, P);
      TTGC_HIP(hipGetLastError());
      std::swap(rz, rzn);
    }
    *iters = it;
    *berr = e;
    return (e <= tol) ? TTG_CHOL_OK : TTG_CHOL_ENOCONV;
  }

  int run(hipStream_t s, const double *A, const double *B, long ldb, double *X, long ldx, double tol, int max_iter,
          int *d_info, int *iters, double *berr) {
    // 1-2. working copy (padding is set later by the factorization's prepare
    //      step), ||A||_inf in the same pass over A, factor
    int rc = convert_and_factor(s, A, 1.0, d_info);
    if (rc) return rc;
    int hinfo = 0;
    TTGC_HIP(hipMemcpyAsync(&hinfo, d_info, sizeof(int), hipMemcpyDeviceToHost, s));
    TTGC_HIP(hipStreamSynchronize(s));
    if (hinfo != 0) {
      *iters = 0;
      *berr = INFINITY;
      return TTG_CHOL_OK;  // caller reads *d_info
    }
    // 3. initial solve (inverses of the diagonal tiles first)
    rc = splan->prepare(s);
    if (rc) return rc;
    rc = solve(s, B, ldb, X, ldx, false);
    if (rc) return rc;
    // 4. refinement; it stops early (ENOCONV) on a NaN backward error or
    // after two steps in a row that did not halve it (a diverging or
    // stagnating refinement would otherwise run all max_iter steps)
    double e = INFINITY, e_prev = INFINITY;
    int it = 0, slow = 0;
    for (;; ++it) {
      rc = residual_fp64(s, N, NB, A, nrhs, X, ldx, B, ldb, r(), N);
      if (rc) return rc;
      rc = berr_launch(s, r(), (long)N, X, ldx, B, ldb);
      if (rc) return rc;
      TTGC_HIP(hipMemcpyAsync(hstats.data(), st(), sizeof(double) * (1 + nrhs), hipMemcpyDeviceToHost, s));
      rc = sync_checked(s);
      if (rc) return rc;
      e = 0.0;
      for (int j = 0; j < nrhs; ++j) e = (hstats[1 + j] > e || std::isnan(hstats[1 + j])) ? hstats[1 + j] : e;
      if (e <= tol || it >= max_iter || std::isnan(e)) break;
      slow = (it > 0 && !(e < 0.5 * e_prev)) ? slow + 1 : 0;
      if (slow >= 2) break;
      e_prev = e;
      rc = solve(s, r(), N, X, ldx, true);
      if (rc) return rc;
    }
    *iters = it;
    *berr = e;
    return (e <= tol) ? TTG_CHOL_OK : TTG_CHOL_ENOCONV;
  }
};

static std::mutex g_mixed_mu;
static std::map<std::tuple<int, int, int, int>, std::unique_ptr<MixedPlanBase>> g_mixed;

void release_mixed_plans() {
  std::lock_guard<std::mutex> lk(g_mixed_mu);
  g_mixed.clear();
}

}  // namespace ttgc

using namespace ttgc;

namespace {

// The cached plan of (N, NB, nrhs, prec), built on first use; the caller
// holds g_mixed_mu.
int mixed_plan(int N, int NB, int nrhs, int prec, MixedPlanBase **out) {
  if (prec != TTG_CHOL_PREC_FP64 && prec != TTG_CHOL_PREC_FP32 && prec != TTG_CHOL_PREC_F16X2 &&
      prec != TTG_CHOL_PREC_F16 && prec != TTG_CHOL_PREC_BF16) {
    set_error("dposv: precision %d not supported", prec);
    return TTG_CHOL_ENOTSUP;
  }
  if (prec >= TTG_CHOL_PREC_F16X2 && NB % 128 != 0) {
    set_error("dposv: FP16/BF16 updates need NB %% 128 == 0 (NB=%d)", NB);
    return TTG_CHOL_ENOTSUP;
  }
  auto key = std::make_tuple(N, NB, nrhs, prec);
  auto &p = g_mixed[key];
  if (!p) {
    int rc;
    if (prec == TTG_CHOL_PREC_FP64) {
      auto *q = new MixedPlan<double>;
      p.reset(q);
      q->N = N;
      q->NB = NB;
      q->nrhs = nrhs;
      rc = q->build();
    } else {
      auto *q = new MixedPlan<float>;
      p.reset(q);
      q->N = N;
      q->NB = NB;
      q->nrhs = nrhs;
      q->lowp = prec == TTG_CHOL_PREC_F16X2 ? kLowF16x2
                : prec == TTG_CHOL_PREC_F16 ? kLowF16
                : prec == TTG_CHOL_PREC_BF16 ? kLowBF16
                                              : 0;
      rc = q->build();
    }
    if (rc) {
      g_mixed.erase(key);
      return rc;
    }
  }
  *out = p.get();
  return TTG_CHOL_OK;
}

bool bad_solve_args(const char *name, int N, int NB, const double *A, int nrhs, const double *B, int ldb,
                    const double *X, int ldx, double tol, int max_iter, const int *d_info, const int *iters,
                    const double *berr) {
  if (N <= 0 || NB <= 0 || NB > TTG_CHOL_MAX_TILE || NB % 32 != 0 || nrhs <= 0 || nrhs > 32 ||
      ldb < std::max(1, N) || ldx < std::max(1, N) || !A || !B || !X || !d_info || !iters || !berr ||
      max_iter < 0 || !(tol > 0.0)) {
    set_error("%s: bad argument (N=%d NB=%d nrhs=%d ldb=%d ldx=%d)", name, N, NB, nrhs, ldb, ldx);
    return true;
  }
  return false;
}

}  // namespace

extern "C" int ttg_chol_dposv_mixed(void *stream, int N, int NB, const double *A, int nrhs, const double *B, int ldb,
                                    double *X, int ldx, int prec, double tol, int max_iter, int *d_info, int *iters,
                                    double *berr) {
  clear_error();
  if (bad_solve_args("dposv_mixed", N, NB, A, nrhs, B, ldb, X, ldx, tol, max_iter, d_info, iters, berr))
    return TTG_CHOL_EINVAL;
  hipStream_t s = as_stream(stream);
  std::lock_guard<std::mutex> lk(g_mixed_mu);
  MixedPlanBase *p = nullptr;
  int rc = mixed_plan(N, NB, nrhs, prec, &p);
  if (rc) return rc;
  if (prec == TTG_CHOL_PREC_FP64)
    return static_cast<MixedPlan<double> *>(p)->run(s, A, B, ldb, X, ldx, tol, max_iter, d_info, iters, berr);
  return static_cast<MixedPlan<float> *>(p)->run(s, A, B, ldb, X, ldx, tol, max_iter, d_info, iters, berr);
}

extern "C" int ttg_chol_dposv_krylov(void *stream, int N, int NB, const double *A, int nrhs, const double *B,
                                     int ldb, double *X, int ldx, int prec, double shift, double tol, int max_iter,
                                     int *d_info, int *iters, double *berr) {
  clear_error();
  if (bad_solve_args("dposv_krylov", N, NB, A, nrhs, B, ldb, X, ldx, tol, max_iter, d_info, iters, berr))
    return TTG_CHOL_EINVAL;
  if (!(shift >= 0.0) || shift > 1.0) {
    set_error("dposv_krylov: shift %g outside [0, 1]", shift);
    return TTG_CHOL_EINVAL;
  }
  hipStream_t s = as_stream(stream);
  std::lock_guard<std::mutex> lk(g_mixed_mu);
  MixedPlanBase *p = nullptr;
  int rc = mixed_plan(N, NB, nrhs, prec, &p);
  if (rc) return rc;
  if (prec == TTG_CHOL_PREC_FP64)
    return static_cast<MixedPlan<double> *>(p)->run_pcg(s, A, B, ldb, X, ldx, tol, max_iter, d_info, shift, iters,
                                                         berr);
  return static_cast<MixedPlan<float> *>(p)->run_pcg(s, A, B, ldb, X, ldx, tol, max_iter, d_info, shift, iters,
                                                      berr);
}

extern "C" int ttg_chol_dposv_mixed_factor(void *stream, int N, int NB, int nrhs, int prec, double *L) {
  clear_error();
  if (!L) {
    set_error("dposv_mixed_factor: L is NULL");
    return TTG_CHOL_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_mixed_mu);
  auto it = g_mixed.find(std::make_tuple(N, NB, nrhs, prec));
  if (it == g_mixed.end() || !it->second) {
    set_error("dposv_mixed_factor: no dposv_mixed call with N=%d NB=%d nrhs=%d prec=%d since the plans were built",
              N, NB, nrhs, prec);
    return TTG_CHOL_ESTATE;
  }
  if (prec == TTG_CHOL_PREC_FP64)
    return static_cast<MixedPlan<double> *>(it->second.get())->factor(as_stream(stream), L);
  return static_cast<MixedPlan<float> *>(it->second.get())->factor(as_stream(stream), L);
}
