// Device-side building blocks of the tiled Cholesky kernels for gfx950
// (CDNA4, 64-lane wavefronts): MFMA wrappers, the LDS-staged double-buffered
// GEMM core shared by the update, TRSM and POTRF kernels, and the
// XCD-aware block remap.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace ttgc {

constexpr int kWave = 64;
constexpr int kXcds = 8;  // MI355X: 8 XCDs, workgroups are dealt round-robin

// Global-memory views of generic pointers.  A load through a generic pointer
// compiles to a FLAT instruction, which counts on both vmcnt and lgkmcnt, so
// every wait for LDS reads (lgkmcnt) also waits for the global loads in
// flight -- a prefetch issued before the LDS reads of a k-slab then completes
// before the slab's first MFMA.  Through these views the accesses are GLOBAL
// instructions, counted on vmcnt only.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gld(const T *p) {
  return (const __attribute__((address_space(1))) T *)p;
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gst(T *p) {
  return (__attribute__((address_space(1))) T *)p;
}

typedef double d2_t __attribute__((ext_vector_type(2)));
typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

// ----------------------------------------------------------------------------
// MFMA 16x16x4 (FP64 and FP32).  Operand layout (64 lanes, l = lane):
//   A fragment 16 x 4:  lane l holds A[i = l % 16][k = l / 16]
//   B fragment 4 x 16:  lane l holds B[k = l / 16][j = l % 16]
//   D 16 x 16:          lane l, element r holds D[i = drow(l / 16, r)][j = l % 16]
// with drow(g, r) = g + 4 r for the FP64 instruction and 4 g + r for FP32
// (measured on gfx950 by the lane-layout test).
// ----------------------------------------------------------------------------
template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
  using acc_t = d4_t;
  static __host__ __device__ constexpr int drow(int g, int r) { return g + 4 * r; }
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

template <>
struct Mfma<float> {
  using acc_t = f4_t;
  static __host__ __device__ constexpr int drow(int g, int r) { return 4 * g + r; }
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

template <typename T>
struct Vec16;  // a 16-byte vector of T
template <>
struct Vec16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
  static constexpr int n = 2;
};
template <>
struct Vec16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  static constexpr int n = 4;
};

// ----------------------------------------------------------------------------
// One-wave factor and inverse of a 32 x 32 SPD block (POTRF and the batched
// solver).
// ----------------------------------------------------------------------------
// Lane l's value of v, for a wave-uniform l.
__device__ __forceinline__ double rdlane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float rdlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// LDS written by some lanes of this wave becomes visible to its other lanes
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 1/sqrt(x) for x > 0: the hardware estimate refined by one Newton step
// (relative error ~ the estimate's squared, below 1 ulp of double/float).
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x * y;
  return y + y * (0.5 - h * y);
}
__device__ __forceinline__ float rsqrt_nr(float x) {
  float y = __builtin_amdgcn_rsqf(x);
  const float h = 0.5f * x * y;
  return y + y * (0.5f - h * y);
}

// Cholesky factor and inverse of the 32 x 32 block in Ds (lower part read)
// by ONE wave, without barriers.  Lane i (mod 32) holds row i of the block in
// registers and factors it left-looking, column by column
//   L(i,j) = (A(i,j) - sum_{k<j} L(i,k) L(j,k)) / L(j,j),
// with row j's finished entries L(j,k) and the pivot taken from lane j by
// readlane (no LDS round trip per column: 26 k instead of 35 k cycles for
// FP32, 37 k instead of 44 k for FP64, measured, bitwise the same factor),
// and 1 / L(j,j) from a refined reciprocal square root (no divisions on the
// chain).  The factor then goes to Ds, and X = inv(L) to Di by 2 x 2 blocks
// of 16: X11 and X22 by forward substitution at once in two lane groups
// (lane c of a group: column c,
//   X(r,c) = -(sum_{c<=k<r} L(r,k) X(k,c)) / L(r,r),  X(c,c) = 1 / L(c,c),
// rows of L as LDS broadcasts), then X21 = -X22 (L21 X11) by two 16 x 16
// MFMA products: 16-step instead of 32-step chains (FP32 26.8 k -> 16.5 k
// cycles for factor and inverse, FP64 37.3 k -> 22.3 k, measured).  Returns
// the 1-based column of the first pivot that is not positive (NaN
// included), 0 if none; on failure Ds and Di are left unspecified.  The
// upper part of Ds must be zero on entry; Di is written whole (zero above
// the diagonal).
template <typename T>
__device__ __forceinline__ int diag_chol_inv_wave(T (*Ds)[33], T (*Di)[33], int lane) {
  constexpr int kB = 32;
  // an opaque copy of the lane index: keeps the compiler from hoisting the
  // 100+ lane masks below out of the caller's loop (and spilling them)
  asm volatile("" : "+v"(lane));
  const int i = lane & 31;
  const bool writer = lane < kB;
  int fail = 0;
  T a[kB];
#pragma unroll
  for (int k = 0; k < kB; ++k) a[k] = (k <= i) ? Ds[i][k] : T(0);
  T rinv = T(0);  // lane i: 1 / L(i,i)
#pragma unroll
  for (int j = 0; j < kB; ++j) {
    T s0 = a[j], s1 = T(0);
#pragma unroll
    for (int k = 0; k < j; k += 2) {
      s0 -= a[k] * rdlane(a[k], j);
      if (k + 1 < j) s1 -= a[k + 1] * rdlane(a[k + 1], j);
    }
    const T s = s0 + s1;
    const T djj = rdlane(s, j);
    fail = (fail == 0 && !(djj > T(0))) ? j + 1 : fail;  // keep going: the caller discards the block
    const T r = rsqrt_nr(djj);
    if (i == j) {
      a[j] = djj * r;  // sqrt(djj)
      rinv = r;
    } else if (i > j) {
      a[j] = s * r;
    }
  }
  if (writer) {
#pragma unroll
    for (int k = 0; k < kB; ++k) Ds[i][k] = a[k];  // zero above the diagonal
  }
  wave_lds_sync();
  if (fail) return fail;
  // X11 (lanes 0-15) and X22 (lanes 32-47): column c of the inverse of
  // L(o + 0..15, o + 0..15), o = 0 / 16; lanes 16-31 zero X12
  {
    const int o = (lane >> 5) << 4, c = lane & 15;
    T x[16];
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) {
      T t0 = T(0), t1 = T(0);
#pragma unroll
      for (int k = 0; k < rr; k += 2) {
        t0 += Ds[o + rr][o + k] * x[k];
        if (k + 1 < rr) t1 += Ds[o + rr][o + k + 1] * x[k + 1];
      }
      const T rl0 = rdlane(rinv, rr), rl1 = rdlane(rinv, 16 + rr);
      const T rl = o ? rl1 : rl0;
      x[rr] = (rr < c) ? T(0) : (rr == c ? rl : -(t0 + t1) * rl);
    }
    if ((lane & 16) == 0) {
#pragma unroll
      for (int rr = 0; rr < 16; ++rr) Di[o + rr][o + c] = x[rr];
    } else if (lane < 32) {
#pragma unroll
      for (int rr = 0; rr < 16; ++rr) Di[rr][16 + c] = T(0);
    }
  }
  wave_lds_sync();
  // L21 X11 into Di(0:16, 16:32) as scratch, then X21 = -X22 (L21 X11) and
  // the scratch back to zero
  const int lr = lane & 15, lk = lane >> 4;
  typename Mfma<T>::acc_t d = {T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int s = 0; s < 4; ++s) d = Mfma<T>::mma(Ds[16 + lr][4 * s + lk], Di[4 * s + lk][lr], d);
#pragma unroll
  for (int r = 0; r < 4; ++r) Di[Mfma<T>::drow(lk, r)][16 + lr] = d[r];
  wave_lds_sync();
  d = typename Mfma<T>::acc_t{T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int s = 0; s < 4; ++s) d = Mfma<T>::mma(Di[16 + lr][16 + 4 * s + lk], Di[4 * s + lk][16 + lr], d);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    Di[16 + Mfma<T>::drow(lk, r)][lr] = -d[r];
    Di[Mfma<T>::drow(lk, r)][16 + lr] = T(0);
  }
  return 0;
}

// Workgroup b of a grid of G is placed on XCD b % 8.  Return a logical index
// such that consecutive logical indices run on the same XCD (and share its
// L2), while the map stays a bijection on [0, G).
__device__ __forceinline__ int xcd_remap(int b, int G) {
  const int q = G / kXcds, r = G % kXcds;
  const int x = b % kXcds, pos = b / kXcds;
  return x * q + (x < r ? x : r) + pos;
}

// s_waitcnt vmcnt(N), expcnt and lgkmcnt not waited for (gfx9 encoding:
// vmcnt bits [3:0] and [15:14], expcnt [6:4], lgkmcnt [11:8])
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// s_waitcnt lgkmcnt(0) only: this wave's LDS reads and writes are done
__device__ __forceinline__ void wait_lgkmcnt0() { __builtin_amdgcn_s_waitcnt(0xc07f); }

// s_barrier without the fence of __syncthreads (which would wait for every
// load in flight); the empty asm keeps the compiler from moving memory
// operations across it
__device__ __forceinline__ void raw_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// LDS-DMA loads (HBM -> LDS, no VGPRs: lane l's 4 or 16 bytes land at the
// wave's LDS base + 4 l or 16 l) written as inline asm, so that the compiler
// does not track them: its tracking makes every LDS read that it cannot
// prove distinct from an LDS-DMA target wait for all LDS DMA in flight,
// which drains a multi-stage ring.  Every wait for these loads is therefore
// explicit (wait_vmcnt; they count on vmcnt in issue order with the other
// vector memory operations).  M0 holds the LDS base; one wait state between
// the write of M0 and the load.
__device__ __forceinline__ void lds_dma_dword(const void *g, const void *l) {
  const unsigned la = (unsigned)(uintptr_t)(const __attribute__((address_space(3))) void *)l;
  asm volatile("s_nop 0\n\tglobal_load_lds_dword %0, off" ::"v"(g), "{m0}"(la) : "memory");
}
__device__ __forceinline__ void lds_dma_dwordx4(const void *g, const void *l) {
  const unsigned la = (unsigned)(uintptr_t)(const __attribute__((address_space(3))) void *)l;
  asm volatile("s_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(g), "{m0}"(la) : "memory");
}
// The same with a wave-uniform 64-bit base in SGPRs and a 32-bit per-lane
// byte offset: the address arithmetic of a stream of such loads is then
// scalar, where per-lane 64-bit products would each cost several VALU
// cycles.  The base is made uniform by readfirstlane (which returns int: the
// halves are widened as unsigned, so a low word >= 2^31 is not sign-extended).
__device__ __forceinline__ unsigned long long uniform_addr(const void *p) {
  const unsigned long long b = (unsigned long long)(uintptr_t)p;
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  return ((unsigned long long)hi << 32) | (unsigned long long)lo;
}
__device__ __forceinline__ void lds_dma_dwordx4_s(const void *sbase, unsigned voff, const void *l) {
  const unsigned la = (unsigned)(uintptr_t)(const __attribute__((address_space(3))) void *)l;
  asm volatile("s_nop 0\n\tglobal_load_lds_dwordx4 %0, %1" ::"v"(voff), "s"(uniform_addr(sbase)), "{m0}"(la)
               : "memory");
}
__device__ __forceinline__ void lds_dma_dword_s(const void *sbase, unsigned voff, const void *l) {
  const unsigned la = (unsigned)(uintptr_t)(const __attribute__((address_space(3))) void *)l;
  asm volatile("s_nop 0\n\tglobal_load_lds_dword %0, %1" ::"v"(voff), "s"(uniform_addr(sbase)), "{m0}"(la)
               : "memory");
}

// Make this workgroup's global stores visible to its own later global loads
// (the POTRF, TRSM and batched kernels re-read data they wrote).  All waves
// of a workgroup run on one CU and share its vector L1, so a workgroup-scope
// fence (wait for the stores) is enough; an agent-scope fence would write
// back this XCD's whole L2 every time.
__device__ __forceinline__ void wg_sync_global() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __syncthreads();
}

// ----------------------------------------------------------------------------
// Stage loader: ROWS x COLS block of a column-major matrix into registers,
// then into an LDS array laid out [COLS][LD] (k-major, rows contiguous).
// FAST: 16-byte loads, no bounds checks (caller guarantees alignment and
// full blocks).  Otherwise scalar predicated loads with zero fill.
// ----------------------------------------------------------------------------
template <typename T, int ROWS, int COLS, int NTH, int LD>
struct StageLoader {
  static constexpr int V = Vec16<T>::n;
  using vec_t = typename Vec16<T>::type;
  static constexpr int VPC = ROWS / V;  // vectors per column
  static_assert(ROWS % V == 0, "rows must be a multiple of the vector width");
  static_assert(NTH % VPC == 0, "threads must cover whole columns");
  static constexpr int CPP = NTH / VPC;  // columns per pass (threads beyond COLS idle)
  static_assert(COLS % CPP == 0 || CPP % COLS == 0, "columns and columns per pass must divide");
  static constexpr int PASSES = COLS >= CPP ? COLS / CPP : 1;
  static constexpr bool PARTIAL = CPP > COLS;

  vec_t r[PASSES];

  template <bool FAST>
  __device__ __forceinline__ void load(const T *__restrict__ base, long ld, int row0, int col0, int rlim,
                                       int clim, int tid) {
    const int v = tid % VPC, c = tid / VPC;
    const int row = row0 + v * V;
    if (PARTIAL && c >= COLS) return;
#pragma unroll
    for (int p = 0; p < PASSES; ++p) {
      const int col = col0 + c + p * CPP;
      if constexpr (FAST) {
        r[p] = *gld(reinterpret_cast<const vec_t *>(base + row + (long)col * ld));
      } else {
#pragma unroll
        for (int e = 0; e < V; ++e)
          r[p][e] = (row + e < rlim && col < clim) ? gld(base)[row + e + (long)col * ld] : T(0);
      }
    }
  }

  __device__ __forceinline__ void store(T (*s)[LD], int tid) const {
    const int v = tid % VPC, c = tid / VPC;
    if (PARTIAL && c >= COLS) return;
#pragma unroll
    for (int p = 0; p < PASSES; ++p) *reinterpret_cast<vec_t *>(&s[c + p * CPP][v * V]) = r[p];
  }
};

// Transposed stage loader: the operand element (j, k) is read from
// base[k + j*ld] (k contiguous), for j in [row0, row0+ROWS), k in
// [col0, col0+COLS), and stored to LDS as s[k][j].
template <typename T, int ROWS, int COLS, int NTH, int LD>
struct StageLoaderT {
  static constexpr int V = Vec16<T>::n;
  using vec_t = typename Vec16<T>::type;
  static constexpr int VPR = COLS / V;  // vectors per row j
  static_assert(COLS % V == 0, "cols must be a multiple of the vector width");
  static_assert(NTH % VPR == 0, "threads must cover whole rows");
  static constexpr int RPP = NTH / VPR;  // rows per pass (threads beyond ROWS idle)
  static_assert(ROWS % RPP == 0 || RPP % ROWS == 0, "rows and rows per pass must divide");
  static constexpr int PASSES = ROWS >= RPP ? ROWS / RPP : 1;
  static constexpr bool PARTIAL = RPP > ROWS;

  vec_t r[PASSES];

  template <bool FAST>
  __device__ __forceinline__ void load(const T *__restrict__ base, long ld, int row0, int col0, int rlim,
                                       int clim, int tid) {
    const int v = tid % VPR, j = tid / VPR;
    const int k = col0 + v * V;
    if (PARTIAL && j >= ROWS) return;
#pragma unroll
    for (int p = 0; p < PASSES; ++p) {
      const int row = row0 + j + p * RPP;
      if constexpr (FAST) {
        r[p] = *gld(reinterpret_cast<const vec_t *>(base + k + (long)row * ld));
      } else {
#pragma unroll
        for (int e = 0; e < V; ++e)
          r[p][e] = (row < rlim && k + e < clim) ? gld(base)[k + e + (long)row * ld] : T(0);
      }
    }
  }

  __device__ __forceinline__ void store(T (*s)[LD], int tid) const {
    const int v = tid % VPR, j = tid / VPR;
    if (PARTIAL && j >= ROWS) return;
#pragma unroll
    for (int p = 0; p < PASSES; ++p)
#pragma unroll
      for (int e = 0; e < V; ++e) s[v * V + e][j + p * RPP] = r[p][e];
  }
};

// ----------------------------------------------------------------------------
// GEMM core: acc (BM x BN block, spread over WM x WN waves) += A * B^T where
// A is rows [i0, i0+BM) and B is rows [j0, j0+BN) of two column-major
// matrices, summed over k in [kbeg, kend) (TB: B's element (j, k) is read
// from B[k + j*ldb] instead; TA: A's element (i, k) from A[k + i*lda]).
// Both operands are staged through
// LDS in BK-wide slabs, double buffered with a register prefetch of the next
// slab while MFMAs consume the current one.
// ----------------------------------------------------------------------------
template <typename T, int BM, int BN, int BK, int WM, int WN, bool TB = false, bool TA = false>
struct GemmCore {
  static constexpr int NW = WM * WN;
  static constexpr int NTH = NW * kWave;
  static constexpr int TM = BM / WM, TN = BN / WN;
  static constexpr int MI = TM / 16, NI = TN / 16;
  static_assert(TM % 16 == 0 && TN % 16 == 0, "wave tile must be a multiple of 16");
  static_assert(BK % 4 == 0, "BK must be a multiple of the MFMA k (4)");
  static constexpr int PAD = 32 / sizeof(T);
  static constexpr int LDA_S = BM + PAD, LDB_S = BN + PAD;
  using acc_t = typename Mfma<T>::acc_t;

  struct Smem {
    T a[2][BK][LDA_S];
    T b[2][BK][LDB_S];
  };

  acc_t acc[MI][NI];

  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
      for (int j = 0; j < NI; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
  }

  __device__ __forceinline__ void compute(const Smem &sm, int buf, int wm, int wn, int lane) {
    const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      T a[MI], b[NI];
#pragma unroll
      for (int i = 0; i < MI; ++i) a[i] = sm.a[buf][ks * 4 + lk][wm * TM + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < NI; ++j) b[j] = sm.b[buf][ks * 4 + lk][wn * TN + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) acc[i][j] = Mfma<T>::mma(a[i], b[j], acc[i][j]);
    }
  }

  // FAST: every slab is a full, 16-byte aligned block (no predication).
  // mlim/nlim: row limits of A and B; kend: column limit of both.
  template <bool FAST>
  __device__ __forceinline__ void run(Smem &sm, const T *__restrict__ A, long lda, const T *__restrict__ B,
                                      long ldb, int i0, int j0, int kbeg, int kend, int mlim, int nlim) {
    run2<FAST, FAST>(sm, A, lda, B, ldb, i0, j0, kbeg, kend, mlim, nlim);
  }

  // FA / FB: the fast (unpredicated, vector) path for the A / B operand.
  template <bool FA, bool FB>
  __device__ __forceinline__ void run2(Smem &sm, const T *__restrict__ A, long lda, const T *__restrict__ B,
                                       long ldb, int i0, int j0, int kbeg, int kend, int mlim, int nlim) {
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const int wm = wave / WN, wn = wave % WN;
    using LoaderA = typename std::conditional<TA, StageLoaderT<T, BM, BK, NTH, LDA_S>,
                                              StageLoader<T, BM, BK, NTH, LDA_S>>::type;
    LoaderA la;
    using LoaderB = typename std::conditional<TB, StageLoaderT<T, BN, BK, NTH, LDB_S>,
                                              StageLoader<T, BN, BK, NTH, LDB_S>>::type;
    LoaderB lb;
    const int nk = (kend - kbeg + BK - 1) / BK;
    if (nk <= 0) return;
    la.template load<FA>(A, lda, i0, kbeg, mlim, kend, tid);
    lb.template load<FB>(B, ldb, j0, kbeg, nlim, kend, tid);
    la.store(sm.a[0], tid);
    lb.store(sm.b[0], tid);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
      const int cur = kt & 1;
      const bool more = kt + 1 < nk;
      if (more) {
        const int k0 = kbeg + (kt + 1) * BK;
        la.template load<FA>(A, lda, i0, k0, mlim, kend, tid);
        lb.template load<FB>(B, ldb, j0, k0, nlim, kend, tid);
      }
      compute(sm, cur, wm, wn, lane);
      if (more) {
        la.store(sm.a[cur ^ 1], tid);
        lb.store(sm.b[cur ^ 1], tid);
      }
      __syncthreads();
    }
  }

  // Row / column (relative to the block origin) of accumulator element r of
  // MFMA tile (i, j) held by this lane.
  static __device__ __forceinline__ int row_of(int wm, int i, int lane, int r) {
    return wm * TM + i * 16 + Mfma<T>::drow(lane >> 4, r);
  }
  static __device__ __forceinline__ int col_of(int wn, int j, int lane) { return wn * TN + j * 16 + (lane & 15); }
};

}  // namespace ttgc
