// Split-FP16 trailing update for the mixed-precision factorization.
//
// After TRSM(k), every row i of every panel tile A_mk (FP32) is scaled by a
// power of two s_i (max_j |A_mk(i,j)| s_i in [2^12, 2^13)) and split into two
// FP16 parts
//   hi = fp16(A s),  lo = fp16(A s - hi),
// stored k-slab-major: [k / 32][row][k % 32], so that the 32-k slab of 128
// consecutive rows a workgroup loads is one contiguous 8 KB block.  The update of tile (m, n) is then
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
template <int MODE>
__global__ __launch_bounds__(256) void split_kernel(const SplitOp *__restrict__ ops, int NB) {
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
      const f4_t x = *gld(reinterpret_cast<const f4_t *>(op.src + r0 + 4 * v + (long)col * NB));
      m0 = fmaxf(m0, fabsf(x[0]));
      m1 = fmaxf(m1, fabsf(x[1]));
      m2 = fmaxf(m2, fabsf(x[2]));
      m3 = fmaxf(m3, fabsf(x[3]));
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
    gst(op.rinv)[r0 + tid] = ldexpf(1.f, e - 13);
  }
  // 2. 64 x 64 blocks: transpose through LDS, split, write row-major
  for (int bc = 0; bc < NB / 64; ++bc) {
    __syncthreads();
    for (int e = tid; e < 64 * 64; e += 256) {
      const int r = e & 63, c = e >> 6;  // coalesced column-major read
      blk[c][r] = gld(op.src)[(r0 + r) + (long)(bc * 64 + c) * NB];
    }
    __syncthreads();
    for (int e = tid; e < 64 * 64; e += 256) {
      const int c = e & 63, r = e >> 6;  // row-major write, k (= c) contiguous
      const float v = blk[c][r] * rsc[r];
      const int k = bc * 64 + c;
      const long o = ((long)(k >> 5) * NB + (r0 + r)) * 32 + (k & 31);  // k-slab-major: [k / 32][row][k % 32]
      if constexpr (MODE == kLowBF16) {
        gst(reinterpret_cast<__bf16 *>(op.hi))[o] = (__bf16)v;
      } else {
        const _Float16 h = (_Float16)v;
        gst(reinterpret_cast<_Float16 *>(op.hi))[o] = h;
        if constexpr (MODE == kLowF16x2) gst(reinterpret_cast<_Float16 *>(op.lo))[o] = (_Float16)(v - (float)h);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// update: 128 x 128 output block per workgroup, 2 x 2 waves of 64 x 64,
// k slabs of 32 staged in LDS (double buffered)
// ---------------------------------------------------------------------------
constexpr int kF16BM = 128, kF16BK = 32;  // LDS rows of 32 halves = 4 chunks of 16 bytes

// 16-byte chunk c of row r (64-byte rows) is stored at chunk c ^ ((r >> 1) & 3):
// the 8 lanes that one LDS cycle serves for a 16-byte read (8 consecutive
// rows, same logical chunk) then cover 8 distinct 16-byte slots of the
// 128-byte bank window (rows alternate halves, the XOR walks the chunks).
__device__ __forceinline__ int swz(int row, int chunk) { return chunk ^ ((row >> 1) & 3); }

template <int NP, int NST = 2>  // parts per operand: 2 (hi, lo) or 1 (hi); LDS stages
struct F16Smem {
  h8_t a[NST][NP][kF16BM][kF16BK / 8];  // [buf][part][row][chunk]
  h8_t b[NST][NP][kF16BM][kF16BK / 8];
};

// the operand parts of one product: A hi, A lo, B hi, B lo and the row factors
struct F16Parts {
  const uint16_t *ah, *al, *bh, *bl;
  const float *ra, *rb;
};

template <int NP>
__device__ __forceinline__ void f16_load(const F16Parts &pp, long ld, int i0, int j0, int k0, int tid,
                                         h8_t (&ra)[2 * NP], h8_t (&rb)[2 * NP]) {
  // 128 rows x 32 k per part and operand = 512 16-byte chunks; 2 per thread
  const int v = tid & 3, r = tid >> 2;  // chunk in row, row 0..63
  const h8_t *ap[2] = {reinterpret_cast<const h8_t *>(pp.ah), reinterpret_cast<const h8_t *>(pp.al)};
  const h8_t *bp[2] = {reinterpret_cast<const h8_t *>(pp.bh), reinterpret_cast<const h8_t *>(pp.bl)};
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const int row = r + 64 * p;
    const long oa = ((long)(k0 >> 5) * ld + (i0 + row)) * 4 + v;  // [k / 32][row][k % 32]
    const long ob = ((long)(k0 >> 5) * ld + (j0 + row)) * 4 + v;
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      ra[NP * p + q] = gld(ap[q])[oa];
      rb[NP * p + q] = gld(bp[q])[ob];
    }
  }
}

template <int NP, int NST>
__device__ __forceinline__ void f16_store(F16Smem<NP, NST> &sm, int buf, int tid, const h8_t (&ra)[2 * NP],
                                          const h8_t (&rb)[2 * NP]) {
  const int v = tid & 3, r = tid >> 2;
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const int row = r + 64 * p, c = swz(row, v);
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      sm.a[buf][q][row][c] = ra[NP * p + q];
      sm.b[buf][q][row][c] = rb[NP * p + q];
    }
  }
}

typedef __bf16 b8_t __attribute__((ext_vector_type(8)));

template <int MODE>
__global__ __launch_bounds__(256, 2) void update_f16x2_kernel(const F16UpdOp *__restrict__ ops, int bpo, int nbn,
                                                           const int *__restrict__ d_abort) {
  constexpr int NP = (MODE == kLowF16x2) ? 2 : 1;
  if (d_abort && *d_abort) return;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int o = b / bpo, rr = b % bpo, bi = rr / nbn, bj = rr % nbn;
  const F16UpdOp op = ops[o];
  if (op.lower && bi < bj) return;
  const int i0 = bi * kF16BM, j0 = bj * kF16BM;
  if (i0 >= op.n || j0 >= op.n) return;
  __shared__ __align__(16) F16Smem<NP, 2> sm;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;
  f4_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};
  // the first slab's operands go out first, then this lane's C values (4
  // consecutive rows of one column per (i, j)), whose latency then hides
  // behind the first slabs instead of delaying them
  // npan products (panels), each op.k deep: slab s of the run is slab
  // s % nk of panel s / nk (a wave-uniform index)
  const int nk = op.k / kF16BK, npan = op.npan > 1 ? op.npan : 1;
  auto part = [&](int q) {
    const long po = q * op.pstride, so = q * op.sstride;
    return F16Parts{op.ahi + po, op.alo + po, op.bhi + po, op.blo + po, op.ra + so, op.rb + so};
  };
  auto load = [&](int s, h8_t(&ra)[2 * NP], h8_t(&rb)[2 * NP]) {
    const int q = s / nk;
    f16_load<NP>(part(q), op.ld, i0, j0, (s - q * nk) * kF16BK, tid, ra, rb);
  };
  h8_t ra0[2 * NP], rb0[2 * NP];
  load(0, ra0, rb0);
  f4_t cv[4][4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float *pc = op.C + (long)(j0 + wn * 64 + j * 16 + lr) * op.ldc + i0 + wm * 64 + 4 * lk;
#pragma unroll
    for (int i = 0; i < 4; ++i) cv[i][j] = *gld(reinterpret_cast<const f4_t *>(pc + i * 16));
  }

  auto compute = [&](int cur) {
    h8_t ah[NP][4], bh[NP][4];
#pragma unroll
    for (int q = 0; q < NP; ++q) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = wm * 64 + i * 16 + lr;
        ah[q][i] = sm.a[cur][q][row][swz(row, lk)];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = wn * 64 + j * 16 + lr;
        bh[q][j] = sm.b[cur][q][row][swz(row, lk)];
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if constexpr (MODE == kLowF16x2) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[1][i], bh[0][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][i], bh[1][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][i], bh[0][j], acc[i][j], 0, 0, 0);
        } else if constexpr (MODE == kLowF16) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][i], bh[0][j], acc[i][j], 0, 0, 0);
        } else {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8_t, ah[0][i]),
                                                               __builtin_bit_cast(b8_t, bh[0][j]), acc[i][j], 0, 0, 0);
        }
      }
  };
  // cv -= acc * (row factor x column factor) of panel q: the products with
  // powers of two are exact, so this is the one rounding of an FP32 update
  auto fold = [&](int q) {
    const F16Parts pq = part(q);
    f4_t rrow[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      rrow[i] = *gld(reinterpret_cast<const f4_t *>(pq.ra + i0 + wm * 64 + i * 16 + 4 * lk));
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float rc = gld(pq.rb)[j0 + wn * 64 + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        cv[i][j] = cv[i][j] - acc[i][j] * (rrow[i] * rc);
        acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};
      }
    }
  };
  const int ns = npan * nk;
  f16_store<NP, 2>(sm, 0, tid, ra0, rb0);
  __syncthreads();
  for (int s = 0; s < ns; ++s) {
    const int cur = s & 1;
    const bool more = s + 1 < ns;
    if (more) load(s + 1, ra0, rb0);
    compute(cur);
    if ((s + 1) % nk == 0 && s + 1 < ns) fold(s / nk);  // a panel done, more to come
    if (more) f16_store<NP, 2>(sm, cur ^ 1, tid, ra0, rb0);
    __syncthreads();
  }
  fold(npan - 1);
  const bool diag = op.lower && bi == bj;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int col = j0 + wn * 64 + j * 16 + lr;
    auto pc = gst(op.C + (long)col * op.ldc + i0 + wm * 64 + 4 * lk);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const f4_t v = cv[i][j];
      const int row = i0 + wm * 64 + i * 16 + 4 * lk;  // rows row .. row + 3
      if (!diag || row >= col) {
        *reinterpret_cast<__attribute__((address_space(1))) f4_t *>(pc + i * 16) = v;
      } else if (row + 3 >= col) {  // the diagonal crosses these 4 rows: lower part only
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (row + r >= col) pc[i * 16 + r] = v[r];
      }
    }
  }
}


hipError_t launch_split(hipStream_t s, const SplitOp *ops, int nops, int NB, int mode) {
  if (nops <= 0) return hipSuccess;
  const dim3 g(nops * (NB / 64)), b(256);
  if (mode == kLowF16x2)
    hipLaunchKernelGGL(split_kernel<kLowF16x2>, g, b, 0, s, ops, NB);
  else if (mode == kLowF16)
    hipLaunchKernelGGL(split_kernel<kLowF16>, g, b, 0, s, ops, NB);
  else
    hipLaunchKernelGGL(split_kernel<kLowBF16>, g, b, 0, s, ops, NB);
  return hipGetLastError();
}

hipError_t launch_update_f16x2(hipStream_t s, const F16UpdOp *ops, int nops, int NB, const int *d_abort, int mode) {
  if (nops <= 0) return hipSuccess;
  const int nbn = NB / kF16BM, bpo = nbn * nbn;
  const long grid = (long)nops * bpo;
  if (grid > 0x7fffffffL) return hipErrorInvalidConfiguration;
  const dim3 g((unsigned)grid), b(256);
  if (mode == kLowF16x2)
    hipLaunchKernelGGL(update_f16x2_kernel<kLowF16x2>, g, b, 0, s, ops, bpo, nbn, d_abort);
  else if (mode == kLowF16)
    hipLaunchKernelGGL(update_f16x2_kernel<kLowF16>, g, b, 0, s, ops, bpo, nbn, d_abort);
  else
    hipLaunchKernelGGL(update_f16x2_kernel<kLowBF16>, g, b, 0, s, ops, bpo, nbn, d_abort);
  return hipGetLastError();
}

}  // namespace ttgc
