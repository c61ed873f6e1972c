// Batched variable-size FP32 SPD factor + solve in ONE persistent launch
// (no reference counterpart: the many-small-matrices case of the solver).
//
// Matrix i is n_i x n_i (n_i <= 1024), column-major, with one right-hand side
// b_i.  The launch has a fixed number of workgroups (a few per CU); each takes
// the next matrix from a device work counter -- the host orders the
// descriptors largest first, so the long factorizations start early and the
// small ones fill in at the end -- and
//   1. factors it in place, left-looking by 32-column blocks: the rows of the
//      block column live in FP32 MFMA accumulators across the waves and are
//      updated by all earlier columns in one streamed GEMM (operands from L2,
//      two k-steps in flight); the 32 x 32 diagonal block is factored and
//      inverted by one wave in registers; the rows below are multiplied by
//      the inverse;
//   2. solves L L^T x = b by blocked forward / backward substitution (the
//      off-diagonal parts as workgroup-wide dot products, the 32 x 32
//      triangles by one wave with readlane broadcasts), x overwriting b.
// info[i] is 0 or the 1-based column of the first non-positive pivot (then
// b_i is left unchanged).  Every wave leaves the work loop when the counter
// passes the batch size.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <vector>

#include "chol_device.hpp"
#include "chol_internal.hpp"

namespace ttgc {

struct BatchDesc {
  float *A;
  float *b;
  int n, lda;
  int idx;  // position in the caller's arrays
  int pad;
};

constexpr int kBB = 32;        // block width
constexpr int kBMaxN = 1024;   // largest matrix of a batch

template <int WAVES>
struct BatchSmem {
  float Ds[kBB][kBB + 1];
  float Di[kBB][kBB + 1];
  float Xs[WAVES][16][kBB + 1];
  float vs[kBMaxN];
  float red[16][kBB];
  int item, fail;
};

// ---------------------------------------------------------------------------
// factorization of one matrix by the whole workgroup
// ---------------------------------------------------------------------------
template <int WAVES>
__device__ int batched_factor(float *__restrict__ A, int lda, int n, BatchSmem<WAVES> &sm, int tid, int lane,
                              int wave) {
  constexpr int RB = kBMaxN / 16 / WAVES;  // 16-row blocks per wave
  typedef __attribute__((address_space(1))) const float gfloat;
  const gfloat *Ag = (const gfloat *)A;
  const int lr = lane & 15, lk = lane >> 4;
  const int nblk = (n + kBB - 1) / kBB;
  int fail = 0;
  for (int jb = 0; jb < nblk; ++jb) {
    const int c0 = jb * kBB, bs = min(kBB, n - c0);
    const int nrb = (n - c0 + 15) / 16;
    const int tact = (nrb - wave + WAVES - 1) / WAVES;  // active row blocks of this wave (uniform)
    f4_t acc[RB][2];
    // MINUS the current block column (lower part), clamped loads then masks
#pragma unroll
    for (int t = 0; t < RB; ++t) {
      const int q = wave + WAVES * t;
      if (t < tact) {
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
          const int col = c0 + jj * 16 + lr;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            acc[t][jj][r] = Ag[min(c0 + q * 16 + Mfma<float>::drow(lk, r), n - 1) + (long)min(col, n - 1) * lda];
        }
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
          const int col = c0 + jj * 16 + lr;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = c0 + q * 16 + Mfma<float>::drow(lk, r);
            acc[t][jj][r] = (row < n && col < n && row >= col) ? -acc[t][jj][r] : 0.f;
          }
        }
      } else {
        acc[t][0] = f4_t{0.f, 0.f, 0.f, 0.f};
        acc[t][1] = f4_t{0.f, 0.f, 0.f, 0.f};
      }
    }
    // acc += L(rows, 0:c0) L(c0 + cols, 0:c0)^T; rows / columns past the edge
    // read row / column n-1 (their outputs are never stored)
    if (c0 > 0) {
      const int ob0 = min(c0 + lr, n - 1), ob1 = min(c0 + 16 + lr, n - 1);
      auto kloop = [&]<int TA>() {
        int oa[TA];
#pragma unroll
        for (int t = 0; t < TA; ++t) oa[t] = min(c0 + (wave + WAVES * t) * 16 + lr, n - 1);
        // operand slots in flight: more when few row blocks leave little
        // MFMA work per k-step to cover the load latency
        constexpr int SL = TA >= 2 ? 2 : 4;
        float bv[SL][2], av[SL][TA];
        auto load = [&](int sl, int k) {
          const gfloat *Ak = Ag + (long)k * lda;
          bv[sl][0] = Ak[ob0];
          bv[sl][1] = Ak[ob1];
#pragma unroll
          for (int t = 0; t < TA; ++t) av[sl][t] = Ak[oa[t]];
        };
        auto mma = [&](int sl) {
#pragma unroll
          for (int t = 0; t < TA; ++t) {
            acc[t][0] = Mfma<float>::mma(av[sl][t], bv[sl][0], acc[t][0]);
            acc[t][1] = Mfma<float>::mma(av[sl][t], bv[sl][1], acc[t][1]);
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
      static_assert(RB <= 8, "at most 8 row blocks per wave");
      switch (tact) {
        case 8: if constexpr (RB >= 8) kloop.template operator()<8>(); break;
        case 7: if constexpr (RB >= 7) kloop.template operator()<7>(); break;
        case 6: if constexpr (RB >= 6) kloop.template operator()<6>(); break;
        case 5: if constexpr (RB >= 5) kloop.template operator()<5>(); break;
        case 4: kloop.template operator()<4>(); break;
        case 3: kloop.template operator()<3>(); break;
        case 2: kloop.template operator()<2>(); break;
        case 1: kloop.template operator()<1>(); break;
        default: break;
      }
    }
#pragma unroll
    for (int t = 0; t < RB; ++t) {
      acc[t][0] = -acc[t][0];
      acc[t][1] = -acc[t][1];
    }
    // diagonal block (row blocks 0 and 1: waves 0 and 1) to LDS, identity padded
    if (wave < 2) {
#pragma unroll
      for (int jj = 0; jj < 2; ++jj)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int rl = wave * 16 + Mfma<float>::drow(lk, r), cl = jj * 16 + lr;
          float v = acc[0][jj][r];
          if (rl >= bs || cl >= bs)
            v = (rl == cl) ? 1.f : 0.f;
          else if (rl < cl)
            v = 0.f;
          sm.Ds[rl][cl] = v;
        }
    }
    __syncthreads();
    if (wave == 0) {
      const int f = diag_chol_inv_wave<float>(sm.Ds, sm.Di, lane);
      if (lane == 0) sm.fail = f;
    }
    __syncthreads();
    if (sm.fail) {
      fail = c0 + sm.fail;
      break;
    }
    for (int e = tid; e < kBB * kBB; e += WAVES * 64) {
      const int r = e & 31, c = e >> 5;
      if (r >= c && r < bs && c < bs) A[c0 + r + (long)(c0 + c) * lda] = sm.Ds[r][c];
    }
    // rows below the diagonal block: X = acc inv(L_D)^T
#pragma unroll
    for (int t = 0; t < RB; ++t) {
      const int q = wave + WAVES * t;
      if (q >= 2 && q < nrb) {
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
          for (int r = 0; r < 4; ++r) sm.Xs[wave][Mfma<float>::drow(lk, r)][jj * 16 + lr] = acc[t][jj][r];
        wave_lds_sync();
        f4_t x0 = {0.f, 0.f, 0.f, 0.f}, x1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < kBB / 4; ++ks) {
          const float av = sm.Xs[wave][lr][ks * 4 + lk];
          x0 = Mfma<float>::mma(av, sm.Di[lr][ks * 4 + lk], x0);
          x1 = Mfma<float>::mma(av, sm.Di[16 + lr][ks * 4 + lk], x1);
        }
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = c0 + q * 16 + Mfma<float>::drow(lk, r);
          if (row < n) {
            if (lr < bs) A[row + (long)(c0 + lr) * lda] = x0[r];
            if (16 + lr < bs) A[row + (long)(c0 + 16 + lr) * lda] = x1[r];
          }
        }
      }
    }
    wg_sync_global();
  }
  return fail;
}

// ---------------------------------------------------------------------------
// x = (L L^T)^{-1} b for one factored matrix, b in sm.vs on entry and exit
// ---------------------------------------------------------------------------
template <int WAVES>
__device__ void batched_solve(const float *__restrict__ A, int lda, int n, BatchSmem<WAVES> &sm, int tid, int lane,
                              int wave) {
  // 16 partial sums per row / column for every workgroup size: the same
  // summation order (and bits) for every launch configuration
  constexpr int NTH = WAVES * 64, G = 16;
  static_assert(NTH >= G * kBB, "too few threads for the partial sums");
  const int nblk = (n + kBB - 1) / kBB;
  // L_jj (lower, identity padded) into sm.Ds; the caller synchronizes
  auto load_diag = [&](int c0, int bs) {
    for (int e = tid; e < kBB * kBB; e += NTH) {
      const int r = e & 31, c = e >> 5;
      sm.Ds[r][c] = (r < bs && c < bs) ? (r >= c ? A[c0 + r + (long)(c0 + c) * lda] : 0.f) : (r == c ? 1.f : 0.f);
    }
  };
  // forward: y_j = L_jj^{-1} (b_j - L_{j,<j} y_{<j})
  for (int jb = 0; jb < nblk; ++jb) {
    const int c0 = jb * kBB, bs = min(kBB, n - c0);
    if (tid < G * kBB) {
      const int r = tid & 31, g = tid >> 5;
      float p = 0.f;
      if (r < bs)
        for (int k = g; k < c0; k += G) p += A[c0 + r + (long)k * lda] * sm.vs[k];
      sm.red[g][r] = p;
    }
    load_diag(c0, bs);
    __syncthreads();
    if (wave == 0) {
      int ln = lane;
      asm volatile("" : "+v"(ln));  // opaque: the per-k lane masks stay inside the loop
      const int r = ln & 31;
      float t = 0.f;
      if (r < bs) {
        t = sm.vs[c0 + r];
#pragma unroll
        for (int g = 0; g < G; ++g) t -= sm.red[g][r];
      }
#pragma unroll
      for (int k = 0; k < kBB; ++k) {
        const float y = rdlane(t, k) / sm.Ds[k][k];
        if (r == k)
          t = y;
        else if (r > k)
          t -= sm.Ds[r][k] * y;
      }
      if (lane < bs) sm.vs[c0 + lane] = t;
    }
    __syncthreads();
  }
  // backward: x_j = L_jj^{-T} (y_j - L_{>j,j}^T x_{>j})
  for (int jb = nblk - 1; jb >= 0; --jb) {
    const int c0 = jb * kBB, bs = min(kBB, n - c0);
    if (tid < G * kBB) {
      const int c = tid / G, g = tid % G;
      float p = 0.f;
      if (c < bs)
        for (int r = c0 + kBB + g; r < n; r += G) p += A[r + (long)(c0 + c) * lda] * sm.vs[r];
      sm.red[g][c] = p;
    }
    load_diag(c0, bs);
    __syncthreads();
    if (wave == 0) {
      int ln = lane;
      asm volatile("" : "+v"(ln));  // opaque: the per-k lane masks stay inside the loop
      const int c = ln & 31;
      float t = 0.f;
      if (c < bs) {
        t = sm.vs[c0 + c];
#pragma unroll
        for (int g = 0; g < G; ++g) t -= sm.red[g][c];
      }
#pragma unroll
      for (int k = kBB - 1; k >= 0; --k) {
        const float x = rdlane(t, k) / sm.Ds[k][k];
        if (c == k)
          t = x;
        else if (c < k)
          t -= sm.Ds[k][c] * x;
      }
      if (lane < bs) sm.vs[c0 + lane] = t;
    }
    __syncthreads();
  }
}

template <int WAVES>
__global__ __launch_bounds__(WAVES * 64) void posv_batched_kernel(const BatchDesc *__restrict__ descs, int count,
                                                                   int *__restrict__ counter, int *__restrict__ info) {
  __shared__ BatchSmem<WAVES> sm;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (;;) {
    if (tid == 0) sm.item = atomicAdd(counter, 1);
    __syncthreads();
    const int it = sm.item;
    if (it >= count) break;  // the same value in every thread: all waves leave
    const BatchDesc d = descs[it];
    const int n = __builtin_amdgcn_readfirstlane(d.n), lda = __builtin_amdgcn_readfirstlane(d.lda);
    const int f = batched_factor<WAVES>(d.A, lda, n, sm, tid, lane, wave);
    if (tid == 0) info[d.idx] = f;
    if (f == 0) {
      for (int i = tid; i < n; i += WAVES * 64) sm.vs[i] = d.b[i];
      __syncthreads();
      batched_solve<WAVES>(d.A, lda, n, sm, tid, lane, wave);
      for (int i = tid; i < n; i += WAVES * 64) d.b[i] = sm.vs[i];
    }
    __syncthreads();  // sm.item and sm.vs are rewritten by the next item
  }
}

// ---------------------------------------------------------------------------
// host side: configurations, descriptor staging, launch
// ---------------------------------------------------------------------------
struct BatchCfg {
  int waves;       // waves per workgroup
  int wg_per_cu;   // persistent workgroups per CU
};
static const BatchCfg kBatchCfgs[] = {{8, 1}, {8, 2}, {16, 1}, {8, 3}};
constexpr int kNumBatchCfgs = sizeof(kBatchCfgs) / sizeof(kBatchCfgs[0]);
constexpr int kDefaultBatchCfg = 1;  // 8 waves, 2 per CU

struct BatchWork {
  DevBuf dev;  // [counter (256 B)][descriptors]
  char *host = nullptr;
  size_t host_bytes = 0;
  hipEvent_t copied = nullptr;
  bool pending = false;
  int cus = 0;
  ~BatchWork() = default;  // lives until the process ends
};
static std::mutex g_batch_mu;
static std::map<hipStream_t, std::unique_ptr<BatchWork>> g_batch;

}  // namespace ttgc

using namespace ttgc;

extern "C" {

int ttg_chol_batched_num_configs(void) { return kNumBatchCfgs; }

int ttg_chol_batched_config(int cfg, int *waves, int *wg_per_cu) {
  clear_error();
  if (cfg < 0 || cfg >= kNumBatchCfgs || !waves || !wg_per_cu) {
    set_error("batched_config: no configuration %d (0..%d)", cfg, kNumBatchCfgs - 1);
    return TTG_CHOL_EINVAL;
  }
  *waves = kBatchCfgs[cfg].waves;
  *wg_per_cu = kBatchCfgs[cfg].wg_per_cu;
  return TTG_CHOL_OK;
}

int ttg_chol_sposv_batched(void *stream, int count, const int *n, float *const *A, const int *lda, float *const *b,
                           int *d_info, int cfg) {
  clear_error();
  if (count < 0 || (count > 0 && (!n || !A || !lda || !b || !d_info))) {
    set_error("sposv_batched: bad argument (count=%d)", count);
    return TTG_CHOL_EINVAL;
  }
  if (cfg < -1 || cfg >= kNumBatchCfgs) {
    set_error("sposv_batched: no configuration %d (-1 or 0..%d)", cfg, kNumBatchCfgs - 1);
    return TTG_CHOL_EINVAL;
  }
  for (int i = 0; i < count; ++i) {
    if (n[i] < 1 || n[i] > kBMaxN || lda[i] < n[i] || !A[i] || !b[i]) {
      set_error("sposv_batched: matrix %d: n=%d (1..%d), lda=%d, A=%p, b=%p", i, n[i], kBMaxN, lda[i],
                (const void *)A[i], (const void *)b[i]);
      return TTG_CHOL_EINVAL;
    }
  }
  if (count == 0) return TTG_CHOL_OK;
  const BatchCfg c = kBatchCfgs[cfg < 0 ? kDefaultBatchCfg : cfg];
  hipStream_t s = as_stream(stream);
  std::lock_guard<std::mutex> lk(g_batch_mu);
  auto &wp = g_batch[s];
  if (!wp) wp.reset(new BatchWork);
  BatchWork &w = *wp;
  if (!w.cus) {
    int dev = 0;
    TTGC_HIP(hipGetDevice(&dev));
    TTGC_HIP(hipDeviceGetAttribute(&w.cus, hipDeviceAttributeMultiprocessorCount, dev));
  }
  // largest first
  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return n[x] > n[y]; });
  const size_t bytes = 256 + sizeof(BatchDesc) * (size_t)count;
  if (w.pending) {
    TTGC_HIP(hipEventSynchronize(w.copied));
    w.pending = false;
  }
  if (!w.copied) TTGC_HIP(hipEventCreateWithFlags(&w.copied, hipEventDisableTiming));
  if (bytes > w.host_bytes) {
    if (w.host) TTGC_HIP(hipHostFree(w.host));
    w.host = nullptr;
    w.host_bytes = 0;
    TTGC_HIP(hipHostMalloc(reinterpret_cast<void **>(&w.host), 2 * bytes, hipHostMallocDefault));
    w.host_bytes = 2 * bytes;
  }
  TTGC_HIP(w.dev.reserve(bytes));
  memset(w.host, 0, 256);  // the work counter starts at 0
  BatchDesc *hd = reinterpret_cast<BatchDesc *>(w.host + 256);
  for (int q = 0; q < count; ++q) {
    const int i = order[q];
    hd[q] = BatchDesc{A[i], b[i], n[i], lda[i], i, 0};
  }
  TTGC_HIP(hipMemcpyAsync(w.dev.ptr, w.host, bytes, hipMemcpyHostToDevice, s));
  TTGC_HIP(hipEventRecord(w.copied, s));
  w.pending = true;
  int *counter = static_cast<int *>(w.dev.ptr);
  const BatchDesc *dd = reinterpret_cast<const BatchDesc *>(static_cast<char *>(w.dev.ptr) + 256);
  const int grid = std::min(count, w.cus * c.wg_per_cu);
  if (c.waves == 16)
    hipLaunchKernelGGL(posv_batched_kernel<16>, dim3(grid), dim3(1024), 0, s, dd, count, counter, d_info);
  else
    hipLaunchKernelGGL(posv_batched_kernel<8>, dim3(grid), dim3(512), 0, s, dd, count, counter, d_info);
  TTGC_HIP(hipGetLastError());
  return TTG_CHOL_OK;
}

}  // extern "C"
