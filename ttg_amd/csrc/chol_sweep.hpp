// Persistent forward + backward sweeps on the tiled factor (internal API).
#pragma once

#include <hip/hip_runtime.h>

namespace ttgc {

constexpr int kSweepRows = 32;     // rows of one work unit
// off-diagonal tiles folded per block step (2 measured no faster: the extra
// slack goes to the whole-block phase, which no longer limits)
constexpr int kSweepDepth = 1;
constexpr int kSweepMaxNB = 512;   // largest tile the kernel's registers cover
constexpr int kSweepMaxRhs = 16;   // right-hand sides per launch (one MFMA column block)
constexpr int kSweepStride = 64;   // ints between two counters (256 bytes)
// counter slots (each kSweepStride ints): the unit counter, the failure
// word, the forward and backward phase-B counts, then per block b the
// forward (2b) and backward (2b+1) phase-A counts
constexpr int kSweepWork = 0, kSweepFail = 1, kSweepFwDone = 2, kSweepBwDone = 3, kSweepBlk0 = 4;

inline int sweep_counter_ints(int nt) { return (kSweepBlk0 + 2 * nt) * kSweepStride; }

// k per wave of the sweep kernel (512 / waves per workgroup)
template <typename T>
constexpr int sweep_wave_k() {
  return sizeof(T) == 4 ? 128 : 64;
}
// no partial tiles and every wave's k range inside a tile
template <typename T>
inline bool sweep_full(int N, int NB) {
  return N % NB == 0 && NB % sweep_wave_k<T>() == 0;
}

template <typename T>
struct SweepArgs {
  const T *tiles;  // lower tile-packed factor
  const T *linv;   // NT inverses of the diagonal tiles (NB x NB, column-major)
  const T *pm;     // P_{m,j} = Linv_m L_{m,m-j} at ((j-1)(NT-1) + m-j) NB^2, j = 1 .. kSweepDepth
  const T *rm;     // R_{i,j} = L_{i+j,i} Linv_i at ((j-1)(NT-1) + i) NB^2
  T *bt;           // Np rows of nrhs values, row i at i * ld: b in, x out
  T *xt;           // Np rows (same layout): y (work)
  int *ctr;        // sweep_counter_ints(NT) ints (zero before the first launch)
  int N, NB, NT, nrhs, ld;  // nrhs <= kSweepMaxRhs columns of a row stride ld
  int full;            // sweep_full<T>(N, NB): no edge masks
  long long timeout;   // wall-clock ticks a workgroup may run before it gives up
};

// Zero the counters (the failure word too if clear_fail), then run both
// sweeps with `grid` workgroups (at least NB / 32 of them must be able to
// be resident at once).
template <typename T>
hipError_t launch_sweep(hipStream_t s, const SweepArgs<T> &args, int grid, bool clear_fail);
// Workgroups of the sweep kernel that fit on one CU.
template <typename T>
hipError_t sweep_occupancy(int *per_cu);

}  // namespace ttgc
