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
// update: 128 x 128 output block per workgroup, 2 x 2 waves of 64 x 64, two
// workgroups per CU.  The 32-k slabs of the A and B panel rows go from HBM
// straight into LDS with gfx950's LDS-DMA loads (global_load_lds_dwordx4:
// each lane's 16 bytes land at the wave's LDS base + 16 lane), so no VGPRs
// stage them and no ds_write runs; two LDS stages, one barrier per stage.
// The loads are inline asm (the compiler does not track them) with the
// slab's base in SGPRs and fixed 32-bit lane offsets: the per-lane 64-bit
// address arithmetic of the plain form, queued behind the wave's MFMAs,
// took most of the time of the k loop.  Every wait is explicit: before the
// barrier of stage t each wave waits (vmcnt) for its own loads of stage t,
// and the barrier then publishes all of them and retires every wave's reads
// of the stage refilled next.
// ---------------------------------------------------------------------------
constexpr int kF16BM = 128, kF16BK = 32;  // LDS rows of 32 halves = 4 chunks of 16 bytes

// 16-byte chunk c of row r (64-byte rows) is stored at chunk c ^ ((r >> 1) & 3):
// the 8 lanes that one LDS cycle serves for a 16-byte read (8 consecutive
// rows, same logical chunk) then cover 8 distinct 16-byte slots of the
// 128-byte bank window (rows alternate halves, the XOR walks the chunks).
__device__ __forceinline__ int swz(int row, int chunk) { return chunk ^ ((row >> 1) & 3); }

typedef __bf16 b8_t __attribute__((ext_vector_type(8)));

// SPS slabs per LDS stage (one barrier per stage): the single-product modes
// hold one part per operand, so two slabs per stage fit the LDS of two
// workgroups per CU and halve the barriers; the split mode keeps one.
template <int MODE, int SPS>
__global__ __launch_bounds__(256, 2) void update_f16x2_kernel(const F16UpdOp *__restrict__ ops, int bpo, int nbn,
                                                           const int *__restrict__ d_abort) {
  constexpr int NP = (MODE == kLowF16x2) ? 2 : 1;  // parts per operand: hi, lo
  constexpr int BM = kF16BM, BN = kF16BM;
  constexpr int A_CH = NP * BM * 4, ST_CH = A_CH + NP * BN * 4;  // 16-byte chunks per slab
  constexpr int NSEG = ST_CH / 64;                                 // 1 KB (16-row) segments per slab
  constexpr int L = NSEG / 4;                                      // LDS-DMA loads per wave per slab
  constexpr int NCV = 16;                                          // C loads per lane
  if (d_abort && *d_abort) return;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int o = b / bpo, rr = b % bpo, bi = rr / nbn, bj = rr % nbn;
  const F16UpdOp op = ops[o];
  if (op.lower && bi < bj) return;
  const int i0 = bi * BM, j0 = bj * BN;
  if (i0 >= op.n || j0 >= op.n) return;
  __shared__ __align__(16) h8_t lds[2][SPS * ST_CH];  // [stage][slab][A parts | B parts][row][chunk]
  __shared__ __align__(16) float scl[2][BM + BN];       // [panel & 1][row factors | column factors]
  // the wave index as a scalar, so that everything derived from it stays in SGPRs
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;
  // npan products (panels), each op.k deep: slab s of the run is slab
  // s % nk of panel s / nk (a wave-uniform index)
  const int nk = op.k / kF16BK, npan = op.npan > 1 ? op.npan : 1;
  const int ns = npan * nk;  // slabs; a multiple of SPS (the launcher checks k)

  // slab s into stage st: per part, wave w loads the 1 KB (16-row) segments
  // w and w + 4 of A and of B; lane l fills physical chunk l & 3 of the
  // segment's row l >> 2 with the logical chunk the swizzle puts there (a
  // permutation within each 64-byte row, so a wave still reads whole 1 KB
  // blocks of the k-slab-major panel).  The lane offsets are fixed for the
  // block; the slab's base is scalar.
  static_assert(L == 4 * NP, "two A and two B segments per wave and part");
  const int row0 = wave * 16 + (lane >> 2);  // segment w + 4 adds 64 rows (4 KB), same swizzle
  const unsigned chk = (unsigned)((lane & 3) ^ ((row0 >> 1) & 3));
  const unsigned voffA = ((unsigned)(i0 + row0) * 4 + chk) * 16, voffB = ((unsigned)(j0 + row0) * 4 + chk) * 16;
  auto issue = [&](int s, int st, int u) {
    const int q = s / nk, ks = s - q * nk;
    const long eo = q * op.pstride + (long)ks * op.ld * kF16BK;  // halves
    const h8_t *l0 = &lds[st][u * ST_CH];
#pragma unroll
    for (int part = 0; part < NP; ++part) {
      const char *ba = reinterpret_cast<const char *>((part ? op.alo : op.ahi) + eo);
      const char *bb = reinterpret_cast<const char *>((part ? op.blo : op.bhi) + eo);
      lds_dma_dwordx4_s(ba, voffA, l0 + part * BM * 4 + wave * 64);
      lds_dma_dwordx4_s(ba + 4096, voffA, l0 + part * BM * 4 + (wave + 4) * 64);
      lds_dma_dwordx4_s(bb, voffB, l0 + A_CH + part * BN * 4 + wave * 64);
      lds_dma_dwordx4_s(bb + 4096, voffB, l0 + A_CH + part * BN * 4 + (wave + 4) * 64);
    }
  };
  // panel q's row and column factors into slot q & 1, by wave 0 (lanes 0-31
  // the BM row factors, lanes 32-63 the BN column factors), one panel ahead
  // of the fold that reads them
  auto issue_scl = [&](int q) {
    if (wave == 0) {
      const long so = q * op.sstride;
      const float *src = lane < 32 ? op.ra + so + i0 + 4 * lane : op.rb + so + j0 + 4 * (lane - 32);
      lds_dma_dwordx4(src, &scl[q & 1][0]);
    }
  };

  f4_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};
  // the first stage's loads (and panel 0's factors) go out first, then this
  // lane's C values (4 consecutive rows of one column per (i, j)), whose
  // latency then hides behind the first slabs instead of delaying them
  issue_scl(0);
#pragma unroll
  for (int u = 0; u < SPS; ++u) issue(u, 0, u);
  f4_t cv[4][4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float *pc = op.C + (long)(j0 + wn * 64 + j * 16 + lr) * op.ldc + i0 + wm * 64 + 4 * lk;
#pragma unroll
    for (int i = 0; i < 4; ++i) cv[i][j] = *gld(reinterpret_cast<const f4_t *>(pc + i * 16));
  }

  auto compute = [&](int st, int u) {
    h8_t ah[NP][4], bh[NP][4];
#pragma unroll
    for (int q = 0; q < NP; ++q) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = wm * 64 + i * 16 + lr;
        ah[q][i] = lds[st][u * ST_CH + q * BM * 4 + row * 4 + swz(row, lk)];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = wn * 64 + j * 16 + lr;
        bh[q][j] = lds[st][u * ST_CH + A_CH + q * BN * 4 + row * 4 + swz(row, lk)];
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
    f4_t rrow[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) rrow[i] = *reinterpret_cast<const f4_t *>(&scl[q & 1][wm * 64 + i * 16 + 4 * lk]);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float rc = scl[q & 1][BM + wn * 64 + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        cv[i][j] = cv[i][j] - acc[i][j] * (rrow[i] * rc);
        acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};
      }
    }
  };
  for (int t = 0; t < ns / SPS; ++t) {
    // this wave's loads of stage t are done once at most the C loads issued
    // after them (stage 0) are outstanding; the barrier makes every wave's
    // visible and retires all reads of stage (t + 1) & 1 (stage t - 1) and,
    // a panel after its fold, of factor slot q & 1
    if (t == 0)
      wait_vmcnt<NCV>();
    else
      wait_vmcnt<0>();
    raw_barrier();
    if ((t + 1) * SPS < ns) {
      if (((t + 1) * SPS) % nk == 0) issue_scl((t + 1) * SPS / nk);
#pragma unroll
      for (int u = 0; u < SPS; ++u) issue((t + 1) * SPS + u, (t + 1) & 1, u);
    }
#pragma unroll
    for (int u = 0; u < SPS; ++u) compute(t & 1, u);
    const int s = (t + 1) * SPS;  // slabs done
    if (s % nk == 0 && s < ns) fold(s / nk - 1);  // a panel done, more to come (nk is a multiple of SPS)
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
  // two slabs per stage need every op's depth k to be a multiple of 64
  // (NB % 128 == 0 in the planned solves)
  if (NB % (2 * kF16BK) != 0 && mode != kLowF16x2) return hipErrorInvalidValue;
  if (mode == kLowF16x2)
    hipLaunchKernelGGL((update_f16x2_kernel<kLowF16x2, 1>), g, b, 0, s, ops, bpo, nbn, d_abort);
  else if (mode == kLowF16)
    hipLaunchKernelGGL((update_f16x2_kernel<kLowF16, 2>), g, b, 0, s, ops, bpo, nbn, d_abort);
  else
    hipLaunchKernelGGL((update_f16x2_kernel<kLowBF16, 2>), g, b, 0, s, ops, bpo, nbn, d_abort);
  return hipGetLastError();
}

}  // namespace ttgc
