// Matrix generators on the device (tile-packed lower layout):
//   PLGSY (examples/potrf/core_plgsy.h:63-150), bit-identical to the
//   reference, and the SPD construction A = B B^T / N + alpha I with integer
//   B (formed exactly on FP64 MFMA) or FP64 uniform B (rounded).  Also the small utility kernels of the
//   tiled drivers (edge-tile padding, info reset, diagonal shift, nrm2).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "chol_device.hpp"
#include "chol_internal.hpp"
#include "chol_kernels.hpp"
#include "chol_tiled.hpp"
#include "chol_tiled_plan.hpp"

namespace ttgc {

__device__ __forceinline__ unsigned long long rnd64_jump(unsigned long long n, unsigned long long seed) {
  // random.h:24-41 (PLASMA LCG skip-ahead by binary powering)
  unsigned long long a_k = 6364136223846793005ULL, c_k = 1ULL, ran = seed;
  for (; n; n >>= 1) {
    if (n & 1ULL) ran = a_k * ran + c_k;
    c_k *= (a_k + 1ULL);
    a_k *= a_k;
  }
  return ran;
}

// 0.5f - ran * 5.4210108624275222e-20f in single precision (core_plgsy.h:86),
// each operation rounded on its own.
__device__ __forceinline__ double plgsy_value(unsigned long long ran) {
#pragma clang fp contract(off)
  const float r = __ull2float_rn(ran);
  const float p = __fmul_rn(r, 5.4210108624275222e-20f);
  return (double)__fsub_rn(0.5f, p);
}

__device__ __forceinline__ double spd_bint(unsigned long long ran) {
  return (double)((int)((((ran >> 32) * 255ULL) >> 32)) - 127);
}

__device__ __forceinline__ void tile_of_index(long t, int nt, int &m, int &n) {
  // inverse of tile_index(m, n) = n*nt - n(n-1)/2 + (m - n)
  int lo = 0, hi = nt - 1;
  while (lo < hi) {  // largest n with start(n) <= t
    const int mid = (lo + hi + 1) / 2;
    const long st = (long)mid * nt - (long)mid * (mid - 1) / 2;
    if (st <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  n = lo;
  m = n + (int)(t - ((long)n * nt - (long)n * (n - 1) / 2));
}

// One wave per tile column: lane l produces rows l, l+64, ... (coalesced).
__global__ __launch_bounds__(256) void plgsy_kernel(int N, int NB, int nt, double bump, unsigned long long seed,
                                                    double *tiles) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64, lane = threadIdx.x & 63;
  const long ncols = (long)nt * (nt + 1) / 2 * NB;
  if (wave >= ncols) return;
  const long t = wave / NB;
  const int jj = wave % NB;
  int m, n;
  tile_of_index(t, nt, m, n);
  const int mb = min(NB, N - m * NB), nb = min(NB, N - n * NB);
  if (jj >= nb) return;
  double *col = tiles + t * (long)NB * NB + (long)jj * NB;
  const long gj = (long)n * NB + jj;
  const unsigned long long N64 = (unsigned long long)N;
  for (int ii = lane; ii < mb; ii += 64) {
    const long gi = (long)m * NB + ii;
    double v;
    if (gi >= gj) {
      v = plgsy_value(rnd64_jump((unsigned long long)gi + (unsigned long long)gj * N64, seed));
      if (gi == gj) v += bump;
    } else {  // upper part of a diagonal tile mirrors the lower part
      v = plgsy_value(rnd64_jump((unsigned long long)gj + (unsigned long long)gi * N64, seed));
    }
    col[ii] = v;
  }
}

// One tile of CORE_dplgsy (core_plgsy.h:63-150) for any tile position:
// entry (gi, gj) takes the LCG value of stream position max + min * gM of its
// global indices (the lower part column-wise, the upper part row-wise, as the
// three branches of the reference generate them) and diagonal tiles add bump.
__global__ __launch_bounds__(256) void plgsy_tile_kernel(double bump, int m, int n, double *A, long lda, long gM,
                                                         long m0, long n0, unsigned long long seed) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64, lane = threadIdx.x & 63;
  if (wave >= n) return;
  const long gj = n0 + wave;
  for (int ii = lane; ii < m; ii += 64) {
    const long gi = m0 + ii;
    const long hi = gi > gj ? gi : gj, lo = gi > gj ? gj : gi;
    double v = plgsy_value(rnd64_jump((unsigned long long)hi + (unsigned long long)lo * (unsigned long long)gM, seed));
    if (m0 == n0 && gi == gj) v += bump;
    A[ii + (long)wave * lda] = v;
  }
}

// The same for a batch of tiles: blockIdx.y selects the tile.
__global__ __launch_bounds__(256) void plgsy_batched_kernel(const PlgsyOp *__restrict__ ops) {
  const PlgsyOp op = ops[blockIdx.y];
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64, lane = threadIdx.x & 63;
  if (wave >= op.n) return;
  const long gj = op.n0 + wave;
  for (int ii = lane; ii < op.m; ii += 64) {
    const long gi = op.m0 + ii;
    const long hi = gi > gj ? gi : gj, lo = gi > gj ? gj : gi;
    double v =
        plgsy_value(rnd64_jump((unsigned long long)hi + (unsigned long long)lo * (unsigned long long)op.gM, op.seed));
    if (op.m0 == op.n0 && gi == gj) v += op.bump;
    op.A[ii + (long)wave * op.lda] = v;
  }
}

// B for the SPD construction (N x N, column-major, doubles).  UNIFORM: FP64
// uniform on [-1, 1), 2 (ran >> 11) 2^-53 - 1 (every step exact); otherwise
// the integer k/127 construction's k.
__device__ __forceinline__ double spd_uval(unsigned long long ran) {
  return 2.0 * ((double)(ran >> 11) * 0x1p-53) - 1.0;
}

template <bool UNIFORM>
__global__ __launch_bounds__(256) void spd_b_kernel(long N, unsigned long long seed, double *B) {
  const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (wave >= N) return;
  const long j = wave;
  // lane l starts at row l; all lanes advance by 64 rows = 64 LCG steps
  unsigned long long ran = rnd64_jump((unsigned long long)(j * N + lane), seed);
  unsigned long long a64 = 6364136223846793005ULL, c64 = 1ULL;
  {  // affine map of 64 steps: (a, c)^64 by squaring
    unsigned long long a = a64, c = c64;
    for (int s = 0; s < 6; ++s) {
      c = c * (a + 1ULL);
      a = a * a;
    }
    a64 = a;
    c64 = c;
  }
  for (long i = lane; i < N; i += 64) {
    B[i + j * N] = UNIFORM ? spd_uval(ran) : spd_bint(ran);
    ran = a64 * ran + c64;
  }
}

__global__ void add_diag_kernel(int N, int NB, int nt, double alpha, double *tiles) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= N) return;
  const int t = (int)(g / NB), r = (int)(g % NB);
  double *d = tiles + tile_index(t, t, nt) * (long)NB * NB;
  d[r + (long)r * NB] = __dadd_rn(d[r + (long)r * NB], alpha);
}

// Reset *d_info and set the padding of the edge tiles (rows/cols >= N of the
// last tile row) to zero, with ones on the padded diagonal, so every tile op
// of the factorization can run on full NB x NB tiles.
template <typename T>
__global__ void prepare_kernel(int N, int NB, int nt, T *tiles, int *d_info) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g == 0 && d_info) *d_info = 0;
  const int rem = N - (nt - 1) * NB;  // valid rows of the last tile row
  if (rem == NB) return;
  // nt tiles in the last tile row, NB x NB each
  const long per = (long)NB * NB;
  if (g >= per * nt) return;
  const int n = (int)(g / per);
  const long e = g % per;
  const int r = (int)(e % NB), c = (int)(e / NB);
  T *t = tiles + tile_index(nt - 1, n, nt) * per;
  if (n < nt - 1) {
    if (r >= rem) t[e] = T(0);
  } else {
    if (r >= rem || c >= rem) t[e] = (r == c) ? T(1) : T(0);
  }
}

__global__ __launch_bounds__(1024) void nrm2_kernel(int n, const double *x, double *out) {
  __shared__ double part[1024];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += x[i] * x[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sqrt(part[0]);
}

// Lane-layout self test of the 16x16x4 MFMA builtins: block m feeds A lane l
// the value l + 1 and B a one-hot 1 in lane m, and stores all 4 outputs of
// every lane to out[(m * 64 + l) * 4 + r].
template <typename T>
__global__ __launch_bounds__(64) void mfma_probe_kernel(double *out) {
  const int m = blockIdx.x, l = threadIdx.x;
  const T a = T(l + 1), b = (l == m) ? T(1) : T(0);
  typename Mfma<T>::acc_t d = {0, 0, 0, 0};
  d = Mfma<T>::mma(a, b, d);
#pragma unroll
  for (int r = 0; r < 4; ++r) out[(m * 64 + l) * 4 + r] = (double)d[r];
}

// FP16 16x16x32 probe: A lane l element e holds l*8 + e + 1; B is one-hot
// at (lane m / 8 ... ) -- block q sets B lane q / 8, element q % 8 to 1.
// out[(q * 64 + l) * 4 + r] = D element r of lane l.
__global__ __launch_bounds__(64) void mfma_probe_f16_kernel(double *out) {
  typedef _Float16 h8 __attribute__((ext_vector_type(8)));
  const int q = blockIdx.x, l = threadIdx.x;
  h8 a, b;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    a[e] = (_Float16)(float)(l * 8 + e + 1);
    b[e] = (_Float16)((l == q / 8 && e == q % 8) ? 1.0f : 0.0f);
  }
  f4_t d = {0, 0, 0, 0};
  d = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, d, 0, 0, 0);
#pragma unroll
  for (int r = 0; r < 4; ++r) out[(q * 64 + l) * 4 + r] = (double)d[r];
}

// FP16 32x32x16 probe: A lane l element e holds l*8 + e + 1; block q sets B
// lane q / 8, element q % 8 to 1.  out[(q * 64 + l) * 16 + r] = D element r
// of lane l.
__global__ __launch_bounds__(64) void mfma_probe_f16_32_kernel(double *out) {
  typedef _Float16 h8 __attribute__((ext_vector_type(8)));
  typedef float f16v __attribute__((ext_vector_type(16)));
  const int q = blockIdx.x, l = threadIdx.x;
  h8 a, b;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    a[e] = (_Float16)(float)(l * 8 + e + 1);
    b[e] = (_Float16)((l == q / 8 && e == q % 8) ? 1.0f : 0.0f);
  }
  f16v d = {};
  d = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, d, 0, 0, 0);
#pragma unroll
  for (int r = 0; r < 16; ++r) out[(q * 64 + l) * 16 + r] = (double)d[r];
}

template <typename T>
hipError_t launch_prepare_t(hipStream_t s, int N, int NB, T *tiles, int *d_info) {
  const int nt = (N + NB - 1) / NB;
  const long work = (long)NB * NB * nt;
  const long blocks = std::max(1L, (work + 255) / 256);
  hipLaunchKernelGGL(prepare_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, N, NB, nt, tiles, d_info);
  return hipGetLastError();
}
template hipError_t launch_prepare_t<double>(hipStream_t, int, int, double *, int *);
template hipError_t launch_prepare_t<float>(hipStream_t, int, int, float *, int *);

hipError_t launch_prepare(hipStream_t s, int N, int NB, double *tiles, int *d_info) {
  return launch_prepare_t<double>(s, N, NB, tiles, d_info);
}

hipError_t launch_plgsy_batched(hipStream_t s, const PlgsyOp *ops, int nops, int max_n) {
  constexpr int kMaxY = 65535;  // grid y limit: launch in chunks of tiles
  for (int i = 0; i < nops; i += kMaxY) {
    const int cnt = std::min(kMaxY, nops - i);
    hipLaunchKernelGGL(plgsy_batched_kernel, dim3((unsigned)((max_n + 3) / 4), (unsigned)cnt), dim3(256), 0, s,
                       ops + i);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_nrm2(hipStream_t s, int n, const double *x, double *r) {
  hipLaunchKernelGGL(nrm2_kernel, dim3(1), dim3(1024), 0, s, n, x, r);
  return hipGetLastError();
}

// A = c B B^T + alpha I into the lower tiles (diagonal tiles whole, both
// triangles), c = 1/(127^2 N) for the integer B (every product and sum exact,
// so any order gives the same bits) or 1/N for the uniform B.
static int spd_generate(hipStream_t s, int N, int NB, double alpha, unsigned long long seed, double *tiles,
                        double *work, bool uniform) {
  if (N == 0) return TTG_CHOL_OK;
  const int nt = (N + NB - 1) / NB;
  if (uniform)
    hipLaunchKernelGGL(spd_b_kernel<true>, dim3((unsigned)(((long)N + 3) / 4)), dim3(256), 0, s, (long)N, seed, work);
  else
    hipLaunchKernelGGL(spd_b_kernel<false>, dim3((unsigned)(((long)N + 3) / 4)), dim3(256), 0, s, (long)N, seed,
                       work);
  TTGC_HIP(hipGetLastError());
  std::vector<UpdOp<double>> fast, slow;
  for (int n = 0; n < nt; ++n)
    for (int m = n; m < nt; ++m) {
      const int mb = std::min(NB, N - m * NB), nb = std::min(NB, N - n * NB);
      UpdOp<double> op{work + (long)m * NB, work + (long)n * NB, tiles + tile_index(m, n, nt) * (long)NB * NB,
                       mb, nb, N, N, N, NB, 0};
      const bool f = update_fast(op, NB);
      (f ? fast : slow).push_back(op);
    }
  const double c = uniform ? 1.0 / (double)N : 1.0 / (127.0 * 127.0 * (double)N);
  DevBuf ops;
  const size_t bytes = sizeof(UpdOp<double>) * (fast.size() + slow.size());
  TTGC_HIP(ops.reserve(bytes));
  auto *d = static_cast<UpdOp<double> *>(ops.ptr);
  if (!fast.empty()) TTGC_HIP(hipMemcpy(d, fast.data(), sizeof(UpdOp<double>) * fast.size(), hipMemcpyHostToDevice));
  if (!slow.empty())
    TTGC_HIP(hipMemcpy(d + fast.size(), slow.data(), sizeof(UpdOp<double>) * slow.size(), hipMemcpyHostToDevice));
  TTGC_HIP(launch_gen_gemm(s, d, (int)fast.size(), NB, NB, c, 0.0, true));
  TTGC_HIP(launch_gen_gemm(s, d + fast.size(), (int)slow.size(), NB, NB, c, 0.0, false));
  hipLaunchKernelGGL(add_diag_kernel, dim3((unsigned)(((long)N + 255) / 256)), dim3(256), 0, s, N, NB, nt, alpha,
                     tiles);
  TTGC_HIP(hipGetLastError());
  // the descriptor buffer is freed when this call returns: wait for its users
  TTGC_HIP(hipStreamSynchronize(s));
  (void)hipFree(ops.ptr);
  return TTG_CHOL_OK;
}

}  // namespace ttgc

using namespace ttgc;

extern "C" {

int ttg_chol_plgsy_tiled(void *stream, int N, int NB, double bump, unsigned long long seed, double *tiles) {
  clear_error();
  if (N < 0 || NB <= 0 || (N > 0 && !tiles)) {
    set_error("plgsy_tiled: bad argument (N=%d NB=%d)", N, NB);
    return TTG_CHOL_EINVAL;
  }
  if (N == 0) return TTG_CHOL_OK;
  const int nt = (N + NB - 1) / NB;
  const long waves = (long)nt * (nt + 1) / 2 * NB;
  const long blocks = (waves + 3) / 4;
  hipLaunchKernelGGL(plgsy_kernel, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), N, NB, nt, bump, seed,
                     tiles);
  TTGC_HIP(hipGetLastError());
  return TTG_CHOL_OK;
}

int ttg_chol_dplgsy_tile(void *stream, double bump, int m, int n, double *A, int lda, int gM, int m0, int n0,
                         unsigned long long seed) {
  clear_error();
  if (m < 0 || n < 0 || lda < std::max(1, m) || m0 < 0 || n0 < 0 || gM < m0 + m || (m > 0 && n > 0 && !A)) {
    set_error("dplgsy_tile: bad argument (m=%d n=%d lda=%d gM=%d m0=%d n0=%d)", m, n, lda, gM, m0, n0);
    return TTG_CHOL_EINVAL;
  }
  if (m == 0 || n == 0) return TTG_CHOL_OK;
  if (group_record_plgsy(as_stream(stream), PlgsyOp{bump, A, (long)lda, (long)gM, (long)m0, (long)n0, m, n, seed}))
    return TTG_CHOL_OK;
  hipLaunchKernelGGL(plgsy_tile_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, as_stream(stream), bump, m, n, A,
                     (long)lda, (long)gM, (long)m0, (long)n0, seed);
  TTGC_HIP(hipGetLastError());
  return TTG_CHOL_OK;
}

int ttg_chol_spd_tiled(void *stream, int N, int NB, double alpha, unsigned long long seed, double *tiles,
                       double *work) {
  clear_error();
  if (N < 0 || NB <= 0 || (N > 0 && (!tiles || !work))) {
    set_error("spd_tiled: bad argument (N=%d NB=%d)", N, NB);
    return TTG_CHOL_EINVAL;
  }
  return spd_generate(as_stream(stream), N, NB, alpha, seed, tiles, work, false);
}

int ttg_chol_spd_uniform_tiled(void *stream, int N, int NB, double alpha, unsigned long long seed, double *tiles,
                               double *work) {
  clear_error();
  if (N < 0 || NB <= 0 || (N > 0 && (!tiles || !work))) {
    set_error("spd_uniform_tiled: bad argument (N=%d NB=%d)", N, NB);
    return TTG_CHOL_EINVAL;
  }
  return spd_generate(as_stream(stream), N, NB, alpha, seed, tiles, work, true);
}

int ttg_chol_selftest_mfma_layout(void *stream, int kind, double *out) {
  clear_error();
  if (!out || kind < 0 || kind > 3) {
    set_error("selftest_mfma_layout: bad argument (kind=%d)", kind);
    return TTG_CHOL_EINVAL;
  }
  if (kind == 0)
    hipLaunchKernelGGL(mfma_probe_kernel<double>, dim3(64), dim3(64), 0, as_stream(stream), out);
  else if (kind == 1)
    hipLaunchKernelGGL(mfma_probe_kernel<float>, dim3(64), dim3(64), 0, as_stream(stream), out);
  else if (kind == 2)
    hipLaunchKernelGGL(mfma_probe_f16_kernel, dim3(512), dim3(64), 0, as_stream(stream), out);
  else
    hipLaunchKernelGGL(mfma_probe_f16_32_kernel, dim3(512), dim3(64), 0, as_stream(stream), out);
  TTGC_HIP(hipGetLastError());
  return TTG_CHOL_OK;
}

int ttg_chol_dnrm2(void *stream, int n, const double *x, double *d_result) {
  clear_error();
  if (n < 0 || !d_result || (n > 0 && !x)) {
    set_error("dnrm2: bad argument (n=%d)", n);
    return TTG_CHOL_EINVAL;
  }
  if (group_defer_nrm2(as_stream(stream), n, x, d_result)) return TTG_CHOL_OK;
  TTGC_HIP(launch_nrm2(as_stream(stream), n, x, d_result));
  return TTG_CHOL_OK;
}

}  // extern "C"
