// Tile kernels of the tiled Cholesky for gfx950.
//
//   potrf_kernel  : one workgroup factors one n x n tile (n <= 512) in place,
//                   blocked by 32: a 32 x 32 diagonal block is factored and
//                   inverted in LDS, the panel below it is multiplied by the
//                   inverse, and the trailing lower triangle is updated with
//                   FP64 MFMA from the LDS-resident panel.  Also emits the
//                   inverses of the 32 x 32 diagonal blocks for the TRSM.
//   trsm_kernel   : B := B L^{-T} by 64-row stripes; each workgroup walks the
//                   32-wide column blocks left to right:  X_j = (B_j -
//                   X_{<j} L_{j,<j}^T) inv(L_jj)^T, the GEMM part on MFMA.
//   update_kernel : batched C := alpha A B^T + beta C (GEMM, or the lower
//                   triangle for SYRK), 128 x 128 blocks per workgroup, LDS
//                   double-buffered, XCD-aware block order.
#include "chol_device.hpp"
#include "chol_kernels.hpp"

namespace ttgc {

constexpr int kPB = 32;                       // POTRF/TRSM inner block
constexpr int kMaxTile = 512;                 // largest single-tile POTRF/TRSM

// ---------------------------------------------------------------------------
// POTRF (one workgroup per tile, FP64 arithmetic, left-looking by 32-column
// blocks).  For block column j, the rows c0..n-1 of that column are held in
// MFMA accumulators spread over the 8 waves (16-row blocks, two waves per
// SIMD, so that one wave's load waits overlap the other's MFMAs); they are updated
// by every earlier block column in one streamed GEMM (operands straight from
// L2, two k-steps in flight), then the 32 x 32 diagonal block is factored and
// inverted by one wave in registers and the rows below are multiplied by the
// inverse.  The waves run at raised priority: the diagonal tiles are the
// critical path of the factorization.
// Float tiles: the finished factor columns go to an FP64 workspace (op.pw,
// n x n) for the later columns to read, and rounded into the tile at once.
// ---------------------------------------------------------------------------
constexpr int kPotrfWaves = 8;                          // two waves per SIMD
constexpr int kRowBlocks = kMaxTile / 16 / kPotrfWaves;  // 16-row blocks per wave

template <typename TS, typename TC>
__global__ __launch_bounds__(kPotrfWaves * 64) void potrf_kernel(const PotrfOpT<TS> *__restrict__ ops, PotrfOpT<TS> one,
                                                    const int *__restrict__ d_abort) {
  using acc_t = typename Mfma<TC>::acc_t;
  if (d_abort && *d_abort) return;
  const PotrfOpT<TS> op = ops ? ops[blockIdx.x] : one;
  __shared__ TC Ds[kPB][kPB + 1];
  __shared__ TC Di[kPB][kPB + 1];
  __shared__ TC Xs[kPotrfWaves][16][kPB + 1];  // per-wave staging of one 16 x 32 block
  __shared__ int fail_s;
  // the diagonal tiles are the critical path of the factorization: issue
  // ahead of co-resident trailing-update waves
  __builtin_amdgcn_s_setprio(3);

  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: lets the row-block tests branch on scalars
  const int lr = lane & 15, lk = lane >> 4;
  const int n = __builtin_amdgcn_readfirstlane(op.n);
  // working matrix in the compute type: the tile itself, or for float
  // tiles factored in FP64 an FP64 workspace that receives the factor columns
  // as they are finished (the float tile is read once, at the
  // initialisation of each block column, and written once, rounded, next to
  // the FP64 copy)
  const long lda = __builtin_amdgcn_readfirstlane(op.lda);
  TC *__restrict__ W;
  long ldw;
  if constexpr (std::is_same_v<TS, TC>) {
    W = op.A;
    ldw = lda;
  } else {
    W = op.pw;
    ldw = n;
  }
  typedef __attribute__((address_space(1))) const TS gts;
  const gts *Ag = (const gts *)op.A;
  auto store = [&](long r, long c, TC v) {
    gst(W)[r + c * ldw] = v;
    if constexpr (!std::is_same_v<TS, TC>) gst(op.A)[r + c * lda] = (TS)v;
  };
  int fail = 0;
  const int nblk = (n + kPB - 1) / kPB;

  for (int jb = 0; jb < nblk; ++jb) {
    const int c0 = jb * kPB, bs = min(kPB, n - c0);
    const int R = n - c0;           // rows of the block column (diagonal block included)
    const int nrb = (R + 15) / 16;  // 16-row blocks
    acc_t acc[kRowBlocks][2];
    typedef __attribute__((address_space(1))) const TC gtc;
    const gtc *Wg = (const gtc *)W;  // tiles and workspace are global memory
    const int tact = (nrb - wave + kPotrfWaves - 1) / kPotrfWaves;  // active row blocks of this wave (uniform)
    // initialise with MINUS the current column block (the GEMM below adds
    // L(rows, k) L(c0+cols, k) and the sign is restored after it); loads from
    // clamped addresses, masked afterwards, so that all of them are in flight
#pragma unroll
    for (int t = 0; t < kRowBlocks; ++t) {
      const int q = wave + kPotrfWaves * t;
      if (t < tact) {
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
          const int col = c0 + jj * 16 + lr;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            acc[t][jj][r] = (TC)Ag[min(c0 + q * 16 + Mfma<TC>::drow(lk, r), n - 1) + (long)min(col, n - 1) * lda];
        }
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
          const int col = c0 + jj * 16 + lr;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = c0 + q * 16 + Mfma<TC>::drow(lk, r);
            acc[t][jj][r] = (row < n && col < n && row >= col) ? -acc[t][jj][r] : TC(0);
          }
        }
      } else {
        acc[t][0] = acc_t{TC(0), TC(0), TC(0), TC(0)};
        acc[t][1] = acc_t{TC(0), TC(0), TC(0), TC(0)};
      }
    }
    // left-looking update with the columns [0, c0), streamed from L2 with the
    // operands of the next 4 columns in flight while 4 are multiplied.  Rows
    // and columns past the tile edge read row / column n-1 instead: they only
    // feed outputs past the edge, which are never stored (the diagonal block
    // pads them with the identity, whose inverse is zero off its block).
    if (c0 > 0) {
      const int ob0 = min(c0 + lr, n - 1), ob1 = min(c0 + 16 + lr, n - 1);
      // straight-line code per count of active row blocks: no branch inside
      // the loop, so the loads of one step stay in flight across the MFMAs of
      // the other
      auto kloop = [&]<int TA>() {
        int oa[TA];
#pragma unroll
        for (int t = 0; t < TA; ++t) oa[t] = min(c0 + (wave + kPotrfWaves * t) * 16 + lr, n - 1);
        // operand slots in flight: more when few row blocks leave little
        // MFMA work per k-step to cover the load latency; 4 for FP32 (one
        // register per operand; 8 would spill)
        constexpr int SL = sizeof(TC) == 4 ? 4 : (TA >= 2 ? 2 : 4);
        TC bv[SL][2], av[SL][TA];
        auto load = [&](int sl, int k) {
          const gtc *Wk = Wg + (long)k * ldw;
          bv[sl][0] = Wk[ob0];
          bv[sl][1] = Wk[ob1];
#pragma unroll
          for (int t = 0; t < TA; ++t) av[sl][t] = Wk[oa[t]];
        };
        auto mma = [&](int sl) {
#pragma unroll
          for (int t = 0; t < TA; ++t) {
            acc[t][0] = Mfma<TC>::mma(av[sl][t], bv[sl][0], acc[t][0]);
            acc[t][1] = Mfma<TC>::mma(av[sl][t], bv[sl][1], acc[t][1]);
          }
        };
#pragma unroll
        for (int sl = 0; sl < SL; ++sl) load(sl, 4 * sl + lk);  // c0 is a multiple of 32 >= 4 SL
        for (int k0 = 0; k0 < c0; k0 += 4 * SL) {
#pragma unroll
          for (int sl = 0; sl < SL; ++sl) {
            mma(sl);
            if (k0 + 4 * (SL + sl) < c0) load(sl, k0 + 4 * (SL + sl) + lk);
          }
        }
      };
      switch (tact) {
        case 4: kloop.template operator()<4>(); break;
        case 3: kloop.template operator()<3>(); break;
        case 2: kloop.template operator()<2>(); break;
        case 1: kloop.template operator()<1>(); break;
        default: break;
      }
    }
#pragma unroll
    for (int t = 0; t < kRowBlocks; ++t) {
      acc[t][0] = -acc[t][0];
      acc[t][1] = -acc[t][1];
    }
    // diagonal block (row blocks 0 and 1, held by waves 0 and 1) to LDS
    if (wave < 2) {
#pragma unroll
      for (int jj = 0; jj < 2; ++jj)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int rl = wave * 16 + Mfma<TC>::drow(lk, r), cl = jj * 16 + lr;
          TC v = acc[0][jj][r];
          if (rl >= bs || cl >= bs) v = (rl == cl) ? TC(1) : TC(0);  // identity padding
          else if (rl < cl) v = TC(0);
          Ds[rl][cl] = v;
        }
    }
    __syncthreads();
    // factor and invert the diagonal block with wave 0; the others wait
    if (wave == 0) {
      const int f = diag_chol_inv_wave(Ds, Di, lane);
      if (lane == 0) fail_s = f;
    }
    __syncthreads();
    if (fail_s) {
      fail = c0 + fail_s;
      break;
    }
    // L_D back to the working matrix; inverse blocks out for the TRSM
    for (int e = tid; e < kPB * kPB; e += kPotrfWaves * 64) {
      const int r = e & 31, c = e >> 5;
      if (r >= c && r < bs && c < bs) store(c0 + r, c0 + c, Ds[r][c]);
      if (op.dinv) gst(op.dinv)[jb * kPB * kPB + e] = (TS)Di[r][c];
    }
    // rows below the diagonal block: X = acc * Di^T (staged through LDS into
    // MFMA A-operand layout, one 16 x 32 block at a time per wave)
#pragma unroll
    for (int t = 0; t < kRowBlocks; ++t) {
      const int q = wave + kPotrfWaves * t;
      if (q >= 2 && q < nrb) {
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
          for (int r = 0; r < 4; ++r) Xs[wave][Mfma<TC>::drow(lk, r)][jj * 16 + lr] = acc[t][jj][r];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        acc_t x0 = {TC(0), TC(0), TC(0), TC(0)}, x1 = {TC(0), TC(0), TC(0), TC(0)};
#pragma unroll
        for (int ks = 0; ks < kPB / 4; ++ks) {
          const TC av = Xs[wave][lr][ks * 4 + lk];
          x0 = Mfma<TC>::mma(av, Di[lr][ks * 4 + lk], x0);
          x1 = Mfma<TC>::mma(av, Di[16 + lr][ks * 4 + lk], x1);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = c0 + q * 16 + Mfma<TC>::drow(lk, r);
          if (row < n) {
            if (lr < bs) store(row, c0 + lr, x0[r]);
            if (16 + lr < bs) store(row, c0 + 16 + lr, x1[r]);
          }
        }
      }
    }
    wg_sync_global();
  }
  if (tid == 0 && op.info) {
    if (fail)
      *op.info = op.info_offset + fail;
    else if (op.write_zero)
      *op.info = 0;
  }
}

// inv(L_jj) of every 32 x 32 diagonal block of an already factored tile.
template <typename TS>
__global__ __launch_bounds__(256) void dinv_kernel(const PotrfOpT<TS> *__restrict__ ops, PotrfOpT<TS> one) {
  const PotrfOpT<TS> op = ops ? ops[blockIdx.x] : one;
  const int n = op.n;
  const long lda = op.lda;
  const int nblk = (n + kPB - 1) / kPB;
  for (int t = threadIdx.x; t < nblk * kPB; t += blockDim.x) {
    const int jb = t / kPB, c = t % kPB, c0 = jb * kPB, bs = min(kPB, n - c0);
    const TS *L = op.A + c0 + (long)c0 * lda;
    double x[kPB];
#pragma unroll
    for (int i = 0; i < kPB; ++i) {
      double v;
      if (c >= bs) {
        v = (i == c) ? 1.0 : 0.0;
      } else if (i < c || i >= bs) {
        v = 0.0;
      } else if (i == c) {
        v = 1.0 / (double)L[c + (long)c * lda];
      } else {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < i; ++k)
          if (k >= c) s += (double)L[i + (long)k * lda] * x[k];
        v = -s / (double)L[i + (long)i * lda];
      }
      x[i] = v;
    }
#pragma unroll
    for (int i = 0; i < kPB; ++i) op.dinv[jb * kPB * kPB + i + c * kPB] = (TS)x[i];
  }
}

// ---------------------------------------------------------------------------
// TRSM  B := B L^{-T}  (TRANS) or  B := B L^{-1}  (!TRANS)
//   TRANS : column blocks left to right,  X_j = (B_j - X_{<j} L_{j,<j}^T) inv(L_jj)^T
//   !TRANS: column blocks right to left,  X_j = (B_j - X_{>j} L_{>j,j}) inv(L_jj)
// ---------------------------------------------------------------------------
template <typename T, bool TRANS>
__global__ __launch_bounds__(256) void trsm_kernel(const TrsmOp<T> *__restrict__ ops, TrsmOp<T> one, int bpo,
                                                   const int *__restrict__ d_abort) {
  if (d_abort && *d_abort) return;
  const int o = blockIdx.x / bpo, s = blockIdx.x % bpo;
  const TrsmOp<T> op = ops ? ops[o] : one;
  const int r0 = s * 64;
  if (r0 >= op.m) return;
  using Core = GemmCore<T, 64, 32, 16, 4, 1, !TRANS>;
  __shared__ __align__(16) typename Core::Smem sm;
  __shared__ T Ts[kPB][64 + 4];
  __shared__ T Dis[kPB][kPB + 1];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  constexpr int V = Vec16<T>::n;
  const bool fast = (op.m % 64 == 0) && (op.n % kPB == 0) && (op.ldb % V == 0) && (op.ldl % V == 0) &&
                    ((reinterpret_cast<uintptr_t>(op.B) & 15) == 0) &&
                    ((reinterpret_cast<uintptr_t>(op.L) & 15) == 0);
  const int nblk = (op.n + kPB - 1) / kPB;
  for (int jj0 = 0; jj0 < nblk; ++jj0) {
    const int j = TRANS ? jj0 : nblk - 1 - jj0;
    const int c0 = j * kPB;
    Core core;
    core.zero();
    if (TRANS) {
      // X_{<j} (cols [0, c0) of this stripe) times rows [c0, c0+32) of L
      if (c0 > 0) {
        if (fast)
          core.template run<true>(sm, op.B, op.ldb, op.L, op.ldl, r0, c0, 0, c0, op.m, op.n);
        else
          core.template run<false>(sm, op.B, op.ldb, op.L, op.ldl, r0, c0, 0, c0, op.m, op.n);
      }
    } else {
      // X_{>j} (cols [c0+32, n)) times L[k, c0+jj] for k > c0+32  (B(jj, k) = L[k + (c0+jj)*ldl])
      const int kb = c0 + kPB;
      if (kb < op.n) {
        if (fast)
          core.template run<true>(sm, op.B, op.ldb, op.L + (long)c0 * op.ldl, op.ldl, r0, 0, kb, op.n, op.m, kPB);
        else
          core.template run<false>(sm, op.B, op.ldb, op.L + (long)c0 * op.ldl, op.ldl, r0, 0, kb, op.n, op.m,
                                   min(kPB, op.n - c0));
      }
    }
    for (int e = tid; e < kPB * kPB; e += 256) Dis[e & 31][e >> 5] = gld(op.dinv)[j * kPB * kPB + e];
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rl = wave * 16 + Mfma<T>::drow(lk, r), cl = jj * 16 + lr;
        const int row = r0 + rl, col = c0 + cl;
        const T bv = (row < op.m && col < op.n) ? gld(op.B)[row + (long)col * op.ldb] : T(0);
        Ts[cl][rl] = bv - core.acc[0][jj][r];
      }
    __syncthreads();
    typename Mfma<T>::acc_t x[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
    for (int ks = 0; ks < kPB / 4; ++ks) {
      const T av = Ts[ks * 4 + lk][wave * 16 + lr];
#pragma unroll
      for (int jj = 0; jj < 2; ++jj) {
        // TRANS: out(r,c) = sum_k T(r,k) Dinv(c,k);  !TRANS: sum_k T(r,k) Dinv(k,c)
        const T bv = TRANS ? Dis[jj * 16 + lr][ks * 4 + lk] : Dis[ks * 4 + lk][jj * 16 + lr];
        x[jj] = Mfma<T>::mma(av, bv, x[jj]);
      }
    }
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = r0 + wave * 16 + Mfma<T>::drow(lk, r), col = c0 + jj * 16 + lr;
        if (row < op.m && col < op.n) gst(op.B)[row + (long)col * op.ldb] = x[jj][r];
      }
    wg_sync_global();
  }
}

// ---------------------------------------------------------------------------
// batched update  C := alpha A B^T + beta C
// ---------------------------------------------------------------------------
// Two block shapes: 128 x 128 for tile updates, 32 x 64 ("skinny") when
// the output has at most 32 rows (the right-hand-side panels of the solves).
template <typename T, bool TB, bool SKINNY, bool TA>
struct UpdCfg;
template <bool TB, bool TA>
struct UpdCfg<double, TB, false, TA> {
  using Core = GemmCore<double, 128, 128, 16, 2, 2, TB, TA>;
};
template <bool TB, bool TA>
struct UpdCfg<float, TB, false, TA> {
  using Core = GemmCore<float, 128, 128, 32, 2, 2, TB, TA>;
};
template <bool TB, bool TA>
struct UpdCfg<double, TB, true, TA> {
  using Core = GemmCore<double, 32, 64, 32, 1, 4, TB, TA>;
};
template <bool TB, bool TA>
struct UpdCfg<float, TB, true, TA> {
  using Core = GemmCore<float, 32, 64, 32, 1, 4, TB, TA>;
};

// FAST applies to both operands for the 128 x 128 shape and to the B operand
// only for the skinny shape (whose A rows are predicated).
template <typename T, bool FAST, bool TB, bool SKINNY, bool TA = false>
__device__ __forceinline__ void update_body(const UpdOp<T> *__restrict__ ops, UpdOp<T> one, int bpo, int nbn,
                                                     T alpha, T beta, const int *__restrict__ d_abort) {
  using Core = typename UpdCfg<T, TB, SKINNY, TA>::Core;
  constexpr int BM = SKINNY ? 32 : 128;
  constexpr int BN = SKINNY ? 64 : 128;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int o = b / bpo, rr = b % bpo, bi = rr / nbn, bj = rr % nbn;
  const UpdOp<T> op = ops ? ops[o] : one;
  if (op.lower && bi < bj) return;
  const int i0 = bi * BM, j0 = bj * BN;
  if (i0 >= op.m || j0 >= op.n) return;
  __shared__ __align__(16) typename Core::Smem sm;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int WN = SKINNY ? 4 : 2;
  const int wm = wave / WN, wn = wave % WN;
  // full float tiles: this lane's C values are loaded before the k loop, so
  // that their latency hides behind it instead of trailing it (measured: a
  // gain for FP32, a loss for FP64, whose 64 doubles per lane crowd the
  // register file)
  constexpr bool PRE = FAST && !SKINNY && sizeof(T) == 4;
  T cpre[PRE ? Core::MI : 1][PRE ? Core::NI : 1][4];
  if constexpr (PRE) {
    if (beta != T(0)) {
#pragma unroll
      for (int i = 0; i < Core::MI; ++i)
#pragma unroll
        for (int j = 0; j < Core::NI; ++j) {
          auto pc = gld(op.C + (long)(j0 + Core::col_of(wn, j, lane)) * op.ldc);
#pragma unroll
          for (int r = 0; r < 4; ++r) cpre[i][j][r] = pc[i0 + Core::row_of(wm, i, lane, r)];
        }
    }
  }
  Core core;
  core.zero();
  core.template run2<FAST && !SKINNY, FAST>(sm, op.A, op.lda, op.B, op.ldb, i0, j0, 0, op.k, op.m, op.n);
  // one epilogue formula on every path (alpha acc rounded, then one fused
  // multiply-add with beta old), so that the unpredicated and the predicated
  // paths, and hence a lone launch and a batched one, round alike
  auto epi = [&](T acc, T old) -> T {
    if constexpr (sizeof(T) == 8)
      return __builtin_fma(beta, old, alpha * acc);
    else
      return __builtin_fmaf(beta, old, alpha * acc);
  };

  const bool diag = op.lower && bi == bj;
  if constexpr (PRE) {
#pragma unroll
    for (int i = 0; i < Core::MI; ++i)
#pragma unroll
      for (int j = 0; j < Core::NI; ++j) {
        const int col = j0 + Core::col_of(wn, j, lane);
        auto pc = gst(op.C + (long)col * op.ldc);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = i0 + Core::row_of(wm, i, lane, r);
          if (!diag || row >= col) {
            pc[row] = (beta == T(0)) ? alpha * core.acc[i][j][r] : epi(core.acc[i][j][r], cpre[i][j][r]);
          }
        }
      }
    return;
  }
  const bool full = FAST && !SKINNY && !diag;
#pragma unroll
  for (int i = 0; i < Core::MI; ++i)
#pragma unroll
    for (int j = 0; j < Core::NI; ++j) {
      const int col = j0 + Core::col_of(wn, j, lane);
      auto pc = gst(op.C + (long)col * op.ldc);
      if (full) {
        if (beta == T(0)) {
#pragma unroll
          for (int r = 0; r < 4; ++r) pc[i0 + Core::row_of(wm, i, lane, r)] = alpha * core.acc[i][j][r];
        } else {
          T old[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) old[r] = pc[i0 + Core::row_of(wm, i, lane, r)];
#pragma unroll
          for (int r = 0; r < 4; ++r) pc[i0 + Core::row_of(wm, i, lane, r)] = epi(core.acc[i][j][r], old[r]);
        }
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = i0 + Core::row_of(wm, i, lane, r);
          if (row < op.m && col < op.n && (!diag || row >= col))
            pc[row] = (beta == T(0)) ? alpha * core.acc[i][j][r] : epi(core.acc[i][j][r], pc[row]);
        }
      }
    }
}

// FP64: two workgroups per CU (the accumulators move from AGPRs into the 256
// VGPRs a wave then has; no spills): one workgroup's barrier and load waits
// overlap the other's MFMAs (measured: the FP64 factorization 28 % faster).
template <typename T, bool FAST, bool TB, bool SKINNY>
__global__ __launch_bounds__(256, sizeof(T) == 8 ? 2 : 1) void update_kernel(const UpdOp<T> *__restrict__ ops, UpdOp<T> one, int bpo, int nbn,
                                                     T alpha, T beta, const int *__restrict__ d_abort) {
  if (d_abort && *d_abort) return;
  update_body<T, FAST, TB, SKINNY>(ops, one, bpo, nbn, alpha, beta, d_abort);
}

// C := alpha A^T op(B) + beta C (A stored k x m): the TRMM, SYRK and GEMM of
// the LAUUM graph (FP64 only)
template <bool FAST, bool TB, bool SKINNY>
__global__ __launch_bounds__(256, 2) void update_ta_kernel(const UpdOp<double> *__restrict__ ops, UpdOp<double> one,
                                                        int bpo, int nbn, double alpha, double beta,
                                                        const int *__restrict__ d_abort) {
  if (d_abort && *d_abort) return;
  update_body<double, FAST, TB, SKINNY, true>(ops, one, bpo, nbn, alpha, beta, d_abort);
}

// The same update, as its own kernel symbol for the matrix generator (so
// profiles keep generation apart from the factorization).
template <typename T, bool FAST>
__global__ __launch_bounds__(256, 2) void gen_gemm_kernel(const UpdOp<T> *__restrict__ ops, UpdOp<T> one, int bpo,
                                                       int nbn, T alpha, T beta) {
  update_body<T, FAST, false, false>(ops, one, bpo, nbn, alpha, beta, nullptr);
}

// ---------------------------------------------------------------------------
// launchers (ops == nullptr: a single op passed by value in `one`)
// ---------------------------------------------------------------------------
template <typename TS>
hipError_t launch_potrf(hipStream_t s, const PotrfOpT<TS> *ops, int nops, const PotrfOpT<TS> &one,
                        const int *d_abort, bool compute_ts) {
  if (nops <= 0) return hipSuccess;
  if constexpr (std::is_same_v<TS, float>) {
    if (compute_ts) {
      hipLaunchKernelGGL((potrf_kernel<float, float>), dim3(nops), dim3(kPotrfWaves * 64), 0, s, ops, one, d_abort);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((potrf_kernel<TS, double>), dim3(nops), dim3(kPotrfWaves * 64), 0, s, ops, one, d_abort);
  return hipGetLastError();
}

template <typename TS>
hipError_t launch_dinv(hipStream_t s, const PotrfOpT<TS> *ops, int nops, const PotrfOpT<TS> &one) {
  if (nops <= 0) return hipSuccess;
  hipLaunchKernelGGL(dinv_kernel<TS>, dim3(nops), dim3(256), 0, s, ops, one);
  return hipGetLastError();
}

template hipError_t launch_potrf<double>(hipStream_t, const PotrfOpT<double> *, int, const PotrfOpT<double> &,
                                         const int *, bool);
template hipError_t launch_potrf<float>(hipStream_t, const PotrfOpT<float> *, int, const PotrfOpT<float> &,
                                        const int *, bool);
template hipError_t launch_dinv<double>(hipStream_t, const PotrfOpT<double> *, int, const PotrfOpT<double> &);
template hipError_t launch_dinv<float>(hipStream_t, const PotrfOpT<float> *, int, const PotrfOpT<float> &);

template <typename T>
hipError_t launch_trsm(hipStream_t s, const TrsmOp<T> *ops, int nops, const TrsmOp<T> &one, int max_m,
                       const int *d_abort, bool trans) {
  if (nops <= 0 || max_m <= 0) return hipSuccess;
  const int bpo = (max_m + 63) / 64;
  if (trans)
    hipLaunchKernelGGL((trsm_kernel<T, true>), dim3(nops * bpo), dim3(256), 0, s, ops, one, bpo, d_abort);
  else
    hipLaunchKernelGGL((trsm_kernel<T, false>), dim3(nops * bpo), dim3(256), 0, s, ops, one, bpo, d_abort);
  return hipGetLastError();
}

template <typename T, bool SKINNY>
static hipError_t launch_update_shape(hipStream_t s, const UpdOp<T> *ops, int nops, const UpdOp<T> &one, int max_m,
                                      int max_n, T alpha, T beta, bool fast, const int *d_abort, bool tb, bool ta) {
  constexpr int BM = SKINNY ? 32 : 128, BN = SKINNY ? 64 : 128;
  const int nbm = (max_m + BM - 1) / BM, nbn = (max_n + BN - 1) / BN;
  const int bpo = nbm * nbn;
  const long grid = (long)nops * bpo;
  if (grid > 0x7fffffffL) return hipErrorInvalidConfiguration;
  const dim3 g((unsigned)grid), b(256);
  if (ta) {
    if constexpr (sizeof(T) == 8) {
      if (tb) {
        if (fast)
          hipLaunchKernelGGL((update_ta_kernel<true, true, SKINNY>), g, b, 0, s, ops, one, bpo, nbn, alpha, beta,
                             d_abort);
        else
          hipLaunchKernelGGL((update_ta_kernel<false, true, SKINNY>), g, b, 0, s, ops, one, bpo, nbn, alpha, beta,
                             d_abort);
      } else {
        if (fast)
          hipLaunchKernelGGL((update_ta_kernel<true, false, SKINNY>), g, b, 0, s, ops, one, bpo, nbn, alpha, beta,
                             d_abort);
        else
          hipLaunchKernelGGL((update_ta_kernel<false, false, SKINNY>), g, b, 0, s, ops, one, bpo, nbn, alpha, beta,
                             d_abort);
      }
      return hipGetLastError();
    } else {
      return hipErrorInvalidValue;  // transposed A is built for FP64 only
    }
  }
  if (tb) {
    if (fast)
      hipLaunchKernelGGL((update_kernel<T, true, true, SKINNY>), g, b, 0, s, ops, one, bpo, nbn, alpha, beta, d_abort);
    else
      hipLaunchKernelGGL((update_kernel<T, false, true, SKINNY>), g, b, 0, s, ops, one, bpo, nbn, alpha, beta,
                         d_abort);
  } else {
    if (fast)
      hipLaunchKernelGGL((update_kernel<T, true, false, SKINNY>), g, b, 0, s, ops, one, bpo, nbn, alpha, beta,
                         d_abort);
    else
      hipLaunchKernelGGL((update_kernel<T, false, false, SKINNY>), g, b, 0, s, ops, one, bpo, nbn, alpha, beta,
                         d_abort);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_update(hipStream_t s, const UpdOp<T> *ops, int nops, const UpdOp<T> &one, int max_m, int max_n,
                         T alpha, T beta, bool fast, const int *d_abort, bool tb, bool ta) {
  if (nops <= 0 || max_m <= 0 || max_n <= 0) return hipSuccess;
  if (max_m <= 32)
    return launch_update_shape<T, true>(s, ops, nops, one, max_m, max_n, alpha, beta, fast, d_abort, tb, ta);
  return launch_update_shape<T, false>(s, ops, nops, one, max_m, max_n, alpha, beta, fast, d_abort, tb, ta);
}

hipError_t launch_gen_gemm(hipStream_t s, const UpdOp<double> *ops, int nops, int max_m, int max_n, double alpha,
                           double beta, bool fast) {
  if (nops <= 0 || max_m <= 0 || max_n <= 0) return hipSuccess;
  const int nbm = (max_m + 127) / 128, nbn = (max_n + 127) / 128;
  const int bpo = nbm * nbn;
  const long grid = (long)nops * bpo;
  if (grid > 0x7fffffffL) return hipErrorInvalidConfiguration;
  UpdOp<double> none{};
  if (fast)
    hipLaunchKernelGGL((gen_gemm_kernel<double, true>), dim3((unsigned)grid), dim3(256), 0, s, ops, none, bpo, nbn,
                       alpha, beta);
  else
    hipLaunchKernelGGL((gen_gemm_kernel<double, false>), dim3((unsigned)grid), dim3(256), 0, s, ops, none, bpo, nbn,
                       alpha, beta);
  return hipGetLastError();
}

template hipError_t launch_trsm<double>(hipStream_t, const TrsmOp<double> *, int, const TrsmOp<double> &, int,
                                        const int *, bool);
template hipError_t launch_trsm<float>(hipStream_t, const TrsmOp<float> *, int, const TrsmOp<float> &, int,
                                       const int *, bool);
template hipError_t launch_update<double>(hipStream_t, const UpdOp<double> *, int, const UpdOp<double> &, int, int,
                                          double, double, bool, const int *, bool, bool);
template hipError_t launch_update<float>(hipStream_t, const UpdOp<float> *, int, const UpdOp<float> &, int, int,
                                         float, float, bool, const int *, bool, bool);

}  // namespace ttgc
