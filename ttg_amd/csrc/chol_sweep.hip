// Both triangular sweeps of a solve on the tiled factor in ONE persistent
// launch (no reference counterpart; the preconditioner solve of the
// refinement loops and the FP64 solve).
//
// The right-hand sides (up to 16 per launch) are held transposed, row i of
// the block vector at i * ld.  With Linv_k = L_kk^{-1} and the products of
// the first D = kSweepDepth off-diagonal tiles of every block row and column
// P_{m,j} = Linv_m L_{m,m-j}, R_{i,j} = L_{i+j,i} Linv_i (j = 1 .. D, from
// the plan's prepare step), the sweeps are
//   forward   y_m = d_m - sum_j P_{m,j} y_{m-j},
//             d_m = Linv_m (b_m - sum_{i<m-D} L_mi y_i)
//   backward  x_i = e_i - sum_j R_{i,j}^T x_{i+j},
//             e_i = Linv_i^T (y_i - sum_{m>i+D} L_mi^T x_m)
// so one block step of the critical chain is a single product of a
// prefetched tile with the block just finished, while d_m (e_i), which
// gathers the whole block, is formed D steps earlier.  Work is cut into
// units of 32 rows of one block row:
//   forward  unit (m, s), m = 0 .. NT-1 (y in xt; c_m in place in bt)
//     phase A   c_m[s] = b_m[s] - sum_{i<m-D} L_mi[s, :] y_i
//     phase B   d_m[s] = Linv_m[s, :] c_m          (all of c_m)
//     phase C   y_m[s] = d_m[s] - sum_j P_{m,j}[s, :] y_{m-j}
//   backward unit (i, s), i = NT-1 .. 0 (x in bt; y updated in place in xt)
//     phase A   c_i[s] = y_i[s] - sum_{m>i+D} L_mi[:, s]^T x_m
//     phase B   e_i[s] = Linv_i[:, s]^T c_i
//     phase C   x_i[s] = e_i[s] - sum_j R_{i,j}[:, s]^T x_{i+j}
// Workgroups take units from a device counter in this order.  A unit reads
// only blocks finished by units with lower indices, except that phase B
// waits for the phase A of its own block's other units; so as long as NB/32
// workgroups are resident, the lowest unfinished unit can always proceed and
// the sweep cannot deadlock.  Completion is counted on the device: per
// block, the phase-A units done (phase B waits for all of them), and per
// direction, the units done (blocks finish in order, so a count of c S means
// blocks 0 .. c-1 are final).  Thread 0 of a workgroup polls, with an
// agent-scope acquire after the count is reached; the storing wave releases
// (agent-scope fence) before it increments a count.  A wait that exceeds the
// plan's time limit, or sees another workgroup's failure, sets the failure
// word and ends the workgroup, so every wave exits even if a count never
// arrives; the host reads the word after the solve (the first launch of the
// next solve clears it).
//
// Inside a unit the k range of every tile is split over the waves; each
// wave accumulates its part over all tiles on MFMA and the parts are summed
// in a fixed order at the end of the phase, so results do not depend on
// timing.  Tiles and vectors are read straight from global memory into
// MFMA operand registers: a tile's loads are issued before the wait for the
// vector it multiplies.
//   rows form (T[s, :] v):   lane l loads 2 consecutive rows
//     r0 + 2 (l % 16) + {0, 1} of column k0 + l / 16 (16 lanes = 128 bytes of
//     one column for FP32), and MFMA e takes element e: its row index q is
//     row r0 + 2 q + e.
//   cols form (T[:, s]^T v):    lane l loads VC consecutive rows
//     k0 + VC (l / 16) + (0 .. VC-1) of column c0 + l % 16 (16 bytes), and
//     MFMA e takes element e with the vector entry of the same k.
#include <hip/hip_runtime.h>

#include "chol_device.hpp"
#include "chol_internal.hpp"
#include "chol_sweep.hpp"
#include "chol_tiled.hpp"

// Timing hook (no code unless a timing harness defines it): event k of
// unit u.
#ifndef TTGC_SWEEP_TRACE
#define TTGC_SWEEP_TRACE(u, k) \
  do {                         \
  } while (0)
#endif

namespace ttgc {

template <typename T>
struct SweepCfg;
template <>
struct SweepCfg<float> {
  static constexpr int NW = 4;  // waves per workgroup (the k split)
};
template <>
struct SweepCfg<double> {
  static constexpr int NW = 8;
};

__global__ void sweep_reset_kernel(int *__restrict__ ctr, int n, int clear_fail) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (clear_fail || i / kSweepStride != kSweepFail) ctr[i] = 0;
}

// A wave-uniform value the compiler must treat as new at this point: keeps
// it from precomputing (and holding) one address per unrolled load.
template <typename P>
__device__ __forceinline__ P opaque(P x) {
  asm volatile("" : "+s"(x));
  return x;
}

// Every wave covers KW = 512 / NW consecutive k of a tile (waves beyond the
// tile idle).  FULL: N % NB == 0 and NB % KW == 0 (no masks).
template <typename T, bool FULL>
__global__ __launch_bounds__(SweepCfg<T>::NW * 64) void sweep_kernel(SweepArgs<T> a) {
  constexpr int NW = SweepCfg<T>::NW;
  constexpr int D = kSweepDepth;
  constexpr int KW = kSweepMaxNB / NW;  // k per wave
  static_assert(KW == sweep_wave_k<T>(), "k per wave");
  constexpr int NKS = KW / 4;           // rows form: k steps per wave
  constexpr int VC = 16 / sizeof(T);    // cols form: vector width along k
  constexpr int MS = 4 * VC;            // cols form: k per macro step
  constexpr int NMS = KW / MS;          // cols form: macro steps per wave
  using acc_t = typename Mfma<T>::acc_t;
  typedef T v2_t __attribute__((ext_vector_type(2)));
  typedef T vc_t __attribute__((ext_vector_type(VC)));

  __shared__ acc_t red[NW - 1][2][64];
  __shared__ int s_unit[2], s_ok[2];

  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lq = lane & 15, lg = lane >> 4;
  const int NB = a.NB, NT = a.NT, nrhs = a.nrhs, ld = a.ld, S = NB / kSweepRows;
  const int nunits = 2 * NT * S;
  const int kw0 = w * KW;          // this wave's k range: kw0 .. kw0 + KW - 1
  const bool wact = kw0 < NB;      // wave-uniform
  const long long deadline = (long long)wall_clock64() + a.timeout;
  int *ctr = a.ctr;
  int *fail = ctr + kSweepFail * kSweepStride;
  int *fw_done = ctr + kSweepFwDone * kSweepStride, *bw_done = ctr + kSweepBwDone * kSweepStride;
  auto blk_cnt = [&](int b, int dir) { return ctr + (kSweepBlk0 + 2 * b + dir) * kSweepStride; };
  auto dim = [&](int b) { return min(NB, a.N - b * NB); };
  auto tile = [&](int m, int n) { return a.tiles + tile_index(m, n, NT) * (long)NB * NB; };

  int par = 0;            // parity of the LDS words thread 0 hands to the workgroup
  int known_fw = 0, known_bw = 0;  // thread 0: the last counts it read
  // Wait until *p >= target.  Thread 0 polls (unless its last read already
  // showed the count) and acquires; the workgroup then passes a barrier.
  // false: the sweep failed (time limit, or another workgroup's failure).
  auto wait_ge = [&](int *p, int target, int &known) -> bool {
    if (tid == 0) {
      int ok = 1;
      if (known < target) {
        for (int spin = 1;; ++spin) {
          known = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (known >= target) break;
          // the failure word and the clock every 32 polls (a poll is one
          // memory round trip; the checks would add another)
          if ((spin & 31) == 0) {
            if (__hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
              ok = 0;
              break;
            }
            if ((long long)wall_clock64() > deadline) {
              __hip_atomic_store(fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              ok = 0;
              break;
            }
          }
          __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
      s_ok[par] = ok;
    }
    __syncthreads();
    const bool r = s_ok[par] != 0;
    par ^= 1;
    return r;
  };
  // wave 0 after storing its outputs: release them, then count the unit
  auto signal = [&](int *p) {
    if (w == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      if (lane == 0) __hip_atomic_fetch_add(p, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  };

  acc_t acc[2];
  auto zero = [&]() { acc[0] = acc[1] = acc_t{0, 0, 0, 0}; };
  // the waves' parts of acc, summed into wave 0 in wave order
  auto reduce = [&]() {
    if (w > 0) {
      red[w - 1][0][lane] = acc[0];
      red[w - 1][1][lane] = acc[1];
    }
    __syncthreads();
    if (w == 0) {
#pragma unroll
      for (int q = 0; q < NW - 1; ++q) {
        acc[0] += red[q][0][lane];
        acc[1] += red[q][1][lane];
      }
    }
  };

  // ---- rows form: acc += t[r0 .. r0+31, :] v, t column-major (ld NB) ----
  // One step: the tile loads go out, then `wait` (the vector's producer),
  // then the vector loads and the MFMAs; the operand registers live only
  // inside the step.  Addresses: a wave-uniform base stepped per k step
  // plus one per-lane offset (the loads take the scalar-base form).
  auto rows_step = [&](const T *t, int r0, int rlim, int klim, const T *v, auto &&wait) -> bool {
    v2_t la[NKS];
    T xr[NKS];
    if (wact) {
      const int row = r0 + 2 * lq;
      const int lo = row + lg * NB;
      const long step = 4L * NB;
      const T *base = t + (long)kw0 * NB;
#pragma unroll
      for (int ks = 0; ks < NKS; ++ks) {
        if constexpr (FULL) {
          la[ks] = *gld(reinterpret_cast<const v2_t *>(base + lo));
        } else {
          const int k = kw0 + 4 * ks + lg;
          v2_t x = {T(0), T(0)};
          if (k < klim) {
            if (row < rlim) x[0] = gld(base)[lo];
            if (row + 1 < rlim) x[1] = gld(base)[lo + 1];
          }
          la[ks] = x;
        }
        base = opaque(base + step);
      }
    }
    if (!wait()) return false;
    if (wact) {
      const int lo = lg * ld + lq;
      const bool jok = lq < nrhs;
      const long step = 4L * ld;
      const T *base = v + (long)kw0 * ld;
#pragma unroll
      for (int ks = 0; ks < NKS; ++ks) {
        const int k = kw0 + 4 * ks + lg;
        xr[ks] = ((FULL || k < klim) && jok) ? gld(base)[lo] : T(0);
        base = opaque(base + step);
      }
#pragma unroll
      for (int ks = 0; ks < NKS; ++ks) {
        acc[0] = Mfma<T>::mma(la[ks][0], xr[ks], acc[0]);
        acc[1] = Mfma<T>::mma(la[ks][1], xr[ks], acc[1]);
      }
    }
    return true;
  };
  // wave 0: the lane's 8 output entries of the rows form (row, j): rows
  // r0 + 2 drow(lg, r) + e, column lq
  T pre[8];  // the entries a phase A subtracts from, loaded when its unit starts
  auto load_pre_rows = [&](const T *out, int r0) {
    if (w != 0 || lq >= nrhs) return;
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
      for (int r = 0; r < 4; ++r) pre[4 * e + r] = gld(out)[(long)(r0 + 2 * Mfma<T>::drow(lg, r) + e) * ld + lq];
  };
  // wave 0: pre := acc (the entries a later product is subtracted from);
  // the same lane positions serve the rows and the cols form
  auto keep_acc = [&]() {
    if (w != 0) return;
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
      for (int r = 0; r < 4; ++r) pre[4 * e + r] = acc[e][r];
  };
  // out(row, j) = pre - acc (sub) or acc
  auto store_rows = [&](T *out, int r0, bool sub) {
    if (w != 0 || lq >= nrhs) return;
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = r0 + 2 * Mfma<T>::drow(lg, r) + e;
        gst(out)[(long)row * ld + lq] = sub ? pre[4 * e + r] - acc[e][r] : acc[e][r];
      }
  };

  // ---- cols form: acc += t[:, c0 .. c0+31]^T v (t has klim valid rows,
  // clim valid columns; v has klim valid rows) ----
  auto cols_step = [&](const T *t, int c0, int klim, int clim, const T *v, auto &&wait) -> bool {
    vc_t lc[2][NMS];
    T xc[NMS][VC];
    if (wact) {
      const T *base = t + kw0;
#pragma unroll
      for (int ms = 0; ms < NMS; ++ms) {
#pragma unroll
        for (int g = 0; g < 2; ++g) {
          const int col = c0 + 16 * g + lq;
          const int lo = VC * lg + col * NB;
          if constexpr (FULL) {
            lc[g][ms] = *gld(reinterpret_cast<const vc_t *>(base + lo));
          } else {
            const int kk = kw0 + MS * ms + VC * lg;
            vc_t x;
#pragma unroll
            for (int e = 0; e < VC; ++e) x[e] = (col < clim && kk + e < klim) ? gld(base)[lo + e] : T(0);
            lc[g][ms] = x;
          }
        }
        base = opaque(base + MS);
      }
    }
    if (!wait()) return false;
    if (wact) {
      const int lo = VC * lg * ld + lq;
      const bool jok = lq < nrhs;
      const long step = (long)MS * ld;
      const T *base = v + (long)kw0 * ld;
#pragma unroll
      for (int ms = 0; ms < NMS; ++ms) {
        const int kk = kw0 + MS * ms + VC * lg;
#pragma unroll
        for (int e = 0; e < VC; ++e)
          xc[ms][e] = ((FULL || kk + e < klim) && jok) ? gld(base)[lo + e * ld] : T(0);
        base = opaque(base + step);
      }
#pragma unroll
      for (int ms = 0; ms < NMS; ++ms) {
#pragma unroll
        for (int e = 0; e < VC; ++e) {
          acc[0] = Mfma<T>::mma(lc[0][ms][e], xc[ms][e], acc[0]);
          acc[1] = Mfma<T>::mma(lc[1][ms][e], xc[ms][e], acc[1]);
        }
      }
    }
    return true;
  };
  // wave 0: output entries of the cols form: rows c0 + 16 g + drow(lg, r)
  auto load_pre_cols = [&](const T *out, int c0) {
    if (w != 0 || lq >= nrhs) return;
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int r = 0; r < 4; ++r) pre[4 * g + r] = gld(out)[(long)(c0 + 16 * g + Mfma<T>::drow(lg, r)) * ld + lq];
  };
  auto store_cols = [&](T *out, int c0, bool sub) {
    if (w != 0 || lq >= nrhs) return;
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = c0 + 16 * g + Mfma<T>::drow(lg, r);
        gst(out)[(long)row * ld + lq] = sub ? pre[4 * g + r] - acc[g][r] : acc[g][r];
      }
  };

  for (int it = 0;; ++it) {
    if (tid == 0) s_unit[it & 1] = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int u = s_unit[it & 1];
    if (u >= nunits) return;
    if (u < NT * S) {
      // ---------------- forward unit (m, s) ----------------
      const int m = u / S, s = u % S, r0 = s * kSweepRows, dm = dim(m);
      const int jmax = min(D, m);  // folded tiles L_{m,m-j}, j = 1 .. jmax
      T *bm = a.bt + (long)m * NB * ld;
      load_pre_rows(bm, r0);  // b_m: written before the launch
      zero();
      for (int i = 0; i < m - D; ++i) {
        const int di = dim(i);
        if (!rows_step(tile(m, i), r0, dm, di, a.xt + (long)i * NB * ld,
                       [&] { return wait_ge(fw_done, (i + 1) * S, known_fw); }))
          return;
      }
      reduce();
      store_rows(bm, r0, true);  // c_m
      signal(blk_cnt(m, 0));
      TTGC_SWEEP_TRACE(u, 1);
      zero();
      int kc = 0;
      if (!rows_step(a.linv + (long)m * NB * NB, r0, dm, dm, bm, [&] {
            const bool r = wait_ge(blk_cnt(m, 0), S, kc);
            TTGC_SWEEP_TRACE(u, 2);
            return r;
          }))
        return;
      reduce();  // d_m
      if (jmax > 0) {
        keep_acc();
        zero();
        // acc = sum_j P_{m,j} y_{m-j}, the farthest first: only the last
        // product (j = 1) waits for the block just finished
        for (int j = jmax; j >= 1; --j) {
          if (!rows_step(a.pm + ((long)(j - 1) * (NT - 1) + (m - j)) * NB * NB, r0, dm, dim(m - j),
                         a.xt + (long)(m - j) * NB * ld, [&] {
                           const bool r = wait_ge(fw_done, (m - j + 1) * S, known_fw);
                           if (j == 1) TTGC_SWEEP_TRACE(u, 0);
                           return r;
                         }))
            return;
        }
        reduce();
      }
      store_rows(a.xt + (long)m * NB * ld, r0, jmax > 0);  // y_m = d_m - sum_j P_{m,j} y_{m-j}
      signal(fw_done);
      TTGC_SWEEP_TRACE(u, 3);
    } else {
      // ---------------- backward unit (i, s) ----------------
      const int v = u - NT * S;
      const int i = NT - 1 - v / S, s = v % S, c0 = s * kSweepRows, di = dim(i);
      const int jmax = min(D, NT - 1 - i);  // folded tiles L_{i+j,i}, j = 1 .. jmax
      T *yi = a.xt + (long)i * NB * ld;
      if (!wait_ge(fw_done, NT * S, known_fw)) return;
      load_pre_cols(yi, c0);  // y_i: final since the forward sweep ended
      zero();
      for (int m = NT - 1; m > i + D; --m) {
        const int dm = dim(m);
        if (!cols_step(tile(m, i), c0, dm, di, a.bt + (long)m * NB * ld,
                       [&] { return wait_ge(bw_done, (NT - m) * S, known_bw); }))
          return;
      }
      reduce();
      store_cols(yi, c0, true);  // c_i
      signal(blk_cnt(i, 1));
      TTGC_SWEEP_TRACE(u, 1);
      zero();
      int kc = 0;
      if (!cols_step(a.linv + (long)i * NB * NB, c0, di, di, yi, [&] {
            const bool r = wait_ge(blk_cnt(i, 1), S, kc);
            TTGC_SWEEP_TRACE(u, 2);
            return r;
          }))
        return;
      reduce();  // e_i
      if (jmax > 0) {
        keep_acc();
        zero();
        for (int j = jmax; j >= 1; --j) {
          if (!cols_step(a.rm + ((long)(j - 1) * (NT - 1) + i) * NB * NB, c0, dim(i + j), di,
                         a.bt + (long)(i + j) * NB * ld, [&] {
                           const bool r = wait_ge(bw_done, (NT - i - j) * S, known_bw);
                           if (j == 1) TTGC_SWEEP_TRACE(u, 0);
                           return r;
                         }))
            return;
        }
        reduce();
      }
      store_cols(a.bt + (long)i * NB * ld, c0, jmax > 0);  // x_i = e_i - sum_j R_{i,j}^T x_{i+j}
      signal(bw_done);
      TTGC_SWEEP_TRACE(u, 3);
    }
  }
}

template <typename T>
hipError_t launch_sweep(hipStream_t s, const SweepArgs<T> &args, int grid, bool clear_fail) {
  const int nctr = sweep_counter_ints(args.NT);
  hipLaunchKernelGGL(sweep_reset_kernel, dim3((nctr + 255) / 256), dim3(256), 0, s, args.ctr, nctr,
                     clear_fail ? 1 : 0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (args.full)
    hipLaunchKernelGGL((sweep_kernel<T, true>), dim3(grid), dim3(SweepCfg<T>::NW * 64), 0, s, args);
  else
    hipLaunchKernelGGL((sweep_kernel<T, false>), dim3(grid), dim3(SweepCfg<T>::NW * 64), 0, s, args);
  return hipGetLastError();
}

template <typename T>
hipError_t sweep_occupancy(int *per_cu) {
  int a = 0, b = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, sweep_kernel<T, true>, SweepCfg<T>::NW * 64, 0);
  if (e != hipSuccess) return e;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, sweep_kernel<T, false>, SweepCfg<T>::NW * 64, 0);
  *per_cu = a < b ? a : b;
  return e;
}

template hipError_t launch_sweep<float>(hipStream_t, const SweepArgs<float> &, int, bool);
template hipError_t launch_sweep<double>(hipStream_t, const SweepArgs<double> &, int, bool);
template hipError_t sweep_occupancy<float>(int *);
template hipError_t sweep_occupancy<double>(int *);

}  // namespace ttgc
