// Split-FP16 trailing update for the mixed-precision factorization.
//
// After TRSM(k), every row i of every panel tile A_mk (FP32) is scaled by a
// power of two s_i (max_j |A_mk(i,j)| s_i in [2^12, 2^13)) and split into two
// FP16 parts
//   hi = fp16(A s),  lo = fp16(A s - hi),
// stored row-major (k contiguous).  The update of tile (m, n) is then
//   C -= diag(1/s_m) (hi_m hi_n^T + hi_m lo_n^T + lo_m hi_n^T) diag(1/s_n)
// with three FP16 MFMAs (v_mfma_f32_16x16x32_f16) accumulating in FP32.  The
// dropped lo lo^T term and the FP16 rounding of lo bound the representation
// error of each operand entry by ~2^-22 of its row's largest entry, close to
// FP32 (2^-24), at a third of the FP16 MFMA rate instead of the FP32 one.
#include <hip/hip_runtime.h>

#include "chol_device.hpp"
#include "chol_f16.hpp"
#include "chol_internal.hpp"

namespace ttgc {

typedef _Float16 h8_t __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// split: one workgroup per 64-row stripe of a panel tile
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void split_kernel(const SplitOp *__restrict__ ops, int NB) {
  __builtin_amdgcn_s_setprio(2);  // panel work: ahead of co-resident update waves
  const int nst = NB / 64;
  const SplitOp op = ops[blockIdx.x / nst];
  const int r0 = (blockIdx.x % nst) * 64;
  __shared__ float red[16][64];
  __shared__ float rsc[64];
  __shared__ float blk[64][65];
  const int tid = threadIdx.x;
  // 1. max |a| of each row of the stripe: thread (v, c) covers rows 4v..4v+3
  //    of the columns c, c + 16, ...
  {
    const int v = tid & 15, c = tid >> 4;
    float m0 = 0.f, m1 = 0.f, m2 = 0.f, m3 = 0.f;
    for (int col = c; col < NB; col += 16) {
      const float4 x = *reinterpret_cast<const float4 *>(op.src + r0 + 4 * v + (long)col * NB);
      m0 = fmaxf(m0, fabsf(x.x));
      m1 = fmaxf(m1, fabsf(x.y));
      m2 = fmaxf(m2, fabsf(x.z));
      m3 = fmaxf(m3, fabsf(x.w));
    }
    red[c][4 * v + 0] = m0;
    red[c][4 * v + 1] = m1;
    red[c][4 * v + 2] = m2;
    red[c][4 * v + 3] = m3;
  }
  __syncthreads();
  if (tid < 64) {
    float mx = 0.f;
#pragma unroll
    for (int c = 0; c < 16; ++c) mx = fmaxf(mx, red[c][tid]);
    // s = 2^(13 - e) with 2^(e-1) <= mx < 2^e, so max|a| s lies in [2^12, 2^13)
    int e = 13;
    if (mx > 0.f && isfinite(mx)) frexpf(mx, &e);
    rsc[tid] = ldexpf(1.f, 13 - e);
    op.rinv[r0 + tid] = ldexpf(1.f, e - 13);
  }
  // 2. 64 x 64 blocks: transpose through LDS, split, write row-major
  for (int bc = 0; bc < NB / 64; ++bc) {
    __syncthreads();
    for (int e = tid; e < 64 * 64; e += 256) {
      const int r = e & 63, c = e >> 6;  // coalesced column-major read
      blk[c][r] = op.src[(r0 + r) + (long)(bc * 64 + c) * NB];
    }
    __syncthreads();
    for (int e = tid; e < 64 * 64; e += 256) {
      const int c = e & 63, r = e >> 6;  // row-major write, k (= c) contiguous
      const float v = blk[c][r] * rsc[r];
      const _Float16 h = (_Float16)v;
      const _Float16 l = (_Float16)(v - (float)h);
      const long o = (long)(r0 + r) * NB + bc * 64 + c;
      reinterpret_cast<_Float16 *>(op.hi)[o] = h;
      reinterpret_cast<_Float16 *>(op.lo)[o] = l;
    }
  }
}

// ---------------------------------------------------------------------------
// update: 128 x 128 output block per workgroup, 2 x 2 waves of 64 x 64,
// k slabs of 32 staged in LDS (double buffered)
// ---------------------------------------------------------------------------
constexpr int kF16BM = 128, kF16BK = 32;  // LDS rows of 32 halves = 4 chunks of 16 bytes

// 16-byte chunk c of row r is stored at chunk c ^ ((r >> 2) & 3): the 16 lanes
// of a quarter-wave that read rows r..r+15 at the same chunk then hit 16
// distinct 16-byte bank groups.
__device__ __forceinline__ int swz(int row, int chunk) { return chunk ^ ((row >> 2) & 3); }

struct F16Smem {
  h8_t a[2][2][kF16BM][kF16BK / 8];  // [buf][hi/lo][row][chunk]
  h8_t b[2][2][kF16BM][kF16BK / 8];
};

__device__ __forceinline__ void f16_load(const F16UpdOp &op, int i0, int j0, int k0, int tid, h8_t (&ra)[4],
                                         h8_t (&rb)[4]) {
  // 128 rows x 32 k x {hi, lo} per operand = 1024 16-byte chunks; 4 per thread
  const int v = tid & 3, r = tid >> 2;  // chunk in row, row 0..63
  const h8_t *ahi = reinterpret_cast<const h8_t *>(op.ahi), *alo = reinterpret_cast<const h8_t *>(op.alo);
  const h8_t *bhi = reinterpret_cast<const h8_t *>(op.bhi), *blo = reinterpret_cast<const h8_t *>(op.blo);
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const int row = r + 64 * p;
    const long oa = ((long)(i0 + row) * op.ld + k0) / 8 + v;
    const long ob = ((long)(j0 + row) * op.ld + k0) / 8 + v;
    ra[2 * p + 0] = ahi[oa];
    ra[2 * p + 1] = alo[oa];
    rb[2 * p + 0] = bhi[ob];
    rb[2 * p + 1] = blo[ob];
  }
}

__device__ __forceinline__ void f16_store(F16Smem &sm, int buf, int tid, const h8_t (&ra)[4], const h8_t (&rb)[4]) {
  const int v = tid & 3, r = tid >> 2;
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const int row = r + 64 * p, c = swz(row, v);
    sm.a[buf][0][row][c] = ra[2 * p + 0];
    sm.a[buf][1][row][c] = ra[2 * p + 1];
    sm.b[buf][0][row][c] = rb[2 * p + 0];
    sm.b[buf][1][row][c] = rb[2 * p + 1];
  }
}

__global__ __launch_bounds__(256) void update_f16x2_kernel(const F16UpdOp *__restrict__ ops, int bpo, int nbn,
                                                           const int *__restrict__ d_abort) {
  if (d_abort && *d_abort) return;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int o = b / bpo, rr = b % bpo, bi = rr / nbn, bj = rr % nbn;
  const F16UpdOp op = ops[o];
  if (op.lower && bi < bj) return;
  const int i0 = bi * kF16BM, j0 = bj * kF16BM;
  if (i0 >= op.n || j0 >= op.n) return;
  __shared__ __align__(16) F16Smem sm;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;
  f4_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};

  h8_t ra[4], rb[4];
  const int nk = op.k / kF16BK;
  f16_load(op, i0, j0, 0, tid, ra, rb);
  f16_store(sm, 0, tid, ra, rb);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) f16_load(op, i0, j0, (kt + 1) * kF16BK, tid, ra, rb);
    h8_t ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = wm * 64 + i * 16 + lr, c = swz(row, lk);
      ah[i] = sm.a[cur][0][row][c];
      al[i] = sm.a[cur][1][row][c];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = wn * 64 + j * 16 + lr, c = swz(row, lk);
      bh[j] = sm.b[cur][0][row][c];
      bl[j] = sm.b[cur][1][row][c];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[i], bh[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bl[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bh[j], acc[i][j], 0, 0, 0);
      }
    if (more) f16_store(sm, cur ^ 1, tid, ra, rb);
    __syncthreads();
  }
  const bool diag = op.lower && bi == bj;
  float rrow[4][4];  // row factors (powers of two: the products below are exact)
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) rrow[i][r] = op.ra[i0 + wm * 64 + i * 16 + Mfma<float>::drow(lk, r)];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int col = j0 + wn * 64 + j * 16 + lr;
    const float rc = op.rb[col];
    float *pc = op.C + (long)col * op.ldc;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = i0 + wm * 64 + i * 16 + Mfma<float>::drow(lk, r);
        if (!diag || row >= col) pc[row] -= acc[i][j][r] * (rrow[i][r] * rc);
      }
  }
}

hipError_t launch_split(hipStream_t s, const SplitOp *ops, int nops, int NB) {
  if (nops <= 0) return hipSuccess;
  hipLaunchKernelGGL(split_kernel, dim3(nops * (NB / 64)), dim3(256), 0, s, ops, NB);
  return hipGetLastError();
}

hipError_t launch_update_f16x2(hipStream_t s, const F16UpdOp *ops, int nops, int NB, const int *d_abort) {
  if (nops <= 0) return hipSuccess;
  const int nbn = NB / kF16BM, bpo = nbn * nbn;
  const long grid = (long)nops * bpo;
  if (grid > 0x7fffffffL) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(update_f16x2_kernel, dim3((unsigned)grid), dim3(256), 0, s, ops, bpo, nbn, d_abort);
  return hipGetLastError();
}

}  // namespace ttgc
