// Tile passes of the per-tile POINV operations (see chol_tileops.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "chol_device.hpp"
#include "chol_tileops.hpp"

namespace ttgc {

// grid (ceil(max_m/32) * ceil(max_n/32), nops): one 32 x 32 block of the
// source of one op per workgroup; transposes go through LDS so that both
// the read and the write are coalesced along columns
__global__ __launch_bounds__(256) void tile_op_kernel(const TopOp *__restrict__ ops, int nbm) {
  const TopOp op = ops[blockIdx.y];
  const int bi = blockIdx.x % nbm, bj = blockIdx.x / nbm;
  const int r0 = bi * 32, c0 = bj * 32;
  if (r0 >= op.m || c0 >= op.n) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  __shared__ double t[32][33];
  switch (op.mode) {
    case kTopTrans:
    case kTopTransLower: {
      // src block (bi, bj) -> dst block (bj, bi)
      for (int k = ty; k < 32; k += 8) {
        const int r = r0 + tx, c = c0 + k;
        t[k][tx] = (r < op.m && c < op.n) ? gld(op.src)[r + (long)c * op.lds] : 0.0;
      }
      __syncthreads();
      for (int k = ty; k < 32; k += 8) {
        const int r = c0 + tx, c = r0 + k;  // dst element (r, c) = src (c, r)
        if (r < op.n && c < op.m) {
          if (op.mode == kTopTrans)
            gst(op.dst)[r + (long)c * op.ldd] = op.alpha * t[tx][k];
          else if (r >= c)
            gst(op.dst)[r + (long)c * op.ldd] = t[tx][k];
        }
      }
      return;
    }
    default:
      for (int k = ty; k < 32; k += 8) {
        const int r = r0 + tx, c = c0 + k;
        if (r >= op.m || c >= op.n) continue;
        const long e = r + (long)c * op.ldd;
        switch (op.mode) {
          case kTopCopy: gst(op.dst)[e] = op.alpha * gld(op.src)[r + (long)c * op.lds]; break;
          case kTopLowerZero: gst(op.dst)[e] = r >= c ? gld(op.src)[r + (long)c * op.lds] : 0.0; break;
          default: gst(op.dst)[e] = r == c ? 1.0 : 0.0; break;
        }
      }
  }
}

// one workgroup per op
__global__ __launch_bounds__(256) void diag_zero_tile_kernel(const DiagZeroOp *__restrict__ ops) {
  const DiagZeroOp op = ops[blockIdx.x];
  __shared__ int best;
  if (threadIdx.x == 0) best = 0x7fffffff;
  __syncthreads();
  int mine = 0x7fffffff;
  for (int i = threadIdx.x; i < op.n; i += blockDim.x)
    if (gld(op.A)[i + (long)i * op.lda] == 0.0 && i < mine) mine = i;
  atomicMin(&best, mine);
  __syncthreads();
  if (threadIdx.x == 0) gst(op.info)[0] = best == 0x7fffffff ? 0 : op.info_offset + best + 1;
}

hipError_t launch_tile_ops(hipStream_t s, const TopOp *ops, int nops, int max_m, int max_n) {
  if (nops <= 0 || max_m <= 0 || max_n <= 0) return hipSuccess;
  const int nbm = (max_m + 31) / 32, nbn = (max_n + 31) / 32;
  for (int o = 0; o < nops; o += 65535) {  // grid y is at most 65535
    hipLaunchKernelGGL(tile_op_kernel, dim3(nbm * nbn, std::min(65535, nops - o)), dim3(256), 0, s, ops + o, nbm);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_diag_zero(hipStream_t s, const DiagZeroOp *ops, int nops) {
  if (nops <= 0) return hipSuccess;
  hipLaunchKernelGGL(diag_zero_tile_kernel, dim3(nops), dim3(256), 0, s, ops);
  return hipGetLastError();
}

}  // namespace ttgc
