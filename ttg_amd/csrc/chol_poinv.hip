// POINV family over the lower tile-packed layout (FP64): TRTRI (L := L^{-1}),
// LAUUM (A := L^T L, lower part) and POTRI (TRTRI then LAUUM), the graphs of
// examples/potrf/trtri_L.h, lauum.h and potri.h.
//
// TRTRI follows the reference's step order.  Step K (trtri_L.h:96-104,
// 178-185, 49-56, 27):
//   TRSMR  A(M,K) := A(M,K) L_KK^{-1}            M > K   (in place; the sign
//          of the reference's alpha = -1 is applied by the GEMM and the
//          final pass)
//   GEMM   A(M,N) -= A(M,K) A(K,N)               M > K, N < K
//   TRSML  Q_N := L_KK^{-1} A(K,N)               N < K   (into a work tile:
//          the reference reads A(K,N) in the GEMMs before its TRSML)
//   final  A(M,K) := -A(M,K), A(K,N) := Q_N, A(K,K) := L_KK^{-1} (lower part)
// L_KK^{-1} of every diagonal tile is formed once up front (TRSM of the
// identity with the 32 x 32 block inverses), since no step changes a
// diagonal tile before its own final pass.  Each step is four batched
// launches on the caller's stream.
//
// LAUUM reads only the original factor (the reference's dispatcher hands
// every task the input tile read-only, lauum.h:190-240), so it is one
// batched GEMM over a transposed copy T of the tiles (T_KN = L_KN^T, the
// diagonal ones L_KK^T with the strict lower part zero):
//   A(M,N) = sum_{K >= M} L_KM^T L_KN = [T_MM T_{M+1,M} ...] [T_MN T_{M+1,N} ...]^T
// where the tiles of one tile column are consecutive in the layout, so both
// operands are plain NB x NB(NT-M) column-major matrices.  The K = M term is
// the reference's TRMM (N < M) or LAUUM (N = M), the others its GEMMs and
// SYRKs (lauum.h:72-80, 37, 172-180, 123-128).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "chol_device.hpp"
#include "chol_internal.hpp"
#include "chol_kernels.hpp"
#include "chol_tiled.hpp"

namespace ttgc {

// one tile-to-tile pass over an NB x NB tile
enum TileMoveMode : int {
  kMoveCopy = 0,        // dst := src
  kMoveCopyLower = 1,   // dst := src on and below the diagonal
  kMoveNegate = 2,      // dst := -src
  kMoveTransLower = 3,  // dst := lower(src)^T  (strict upper of src read as zero)
  kMoveTrans = 4,       // dst := src^T
  kMoveIdentity = 5,    // dst := I
};

struct TileMove {
  const double *src;
  double *dst;
  int mode;
};

// grid (NB/32 * NB/32, nops): one 32 x 32 block of one tile per workgroup
__global__ __launch_bounds__(256) void tile_move_kernel(const TileMove *__restrict__ ops, int NB,
                                                        const int *__restrict__ d_abort) {
  if (d_abort && *d_abort) return;
  const TileMove op = ops[blockIdx.y];
  const int nb = NB / 32;
  const int bi = blockIdx.x % nb, bj = blockIdx.x / nb;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  __shared__ double t[32][33];
  if (op.mode == kMoveTransLower || op.mode == kMoveTrans) {
    // dst block (bi, bj) is the transpose of src block (bj, bi)
    for (int k = ty; k < 32; k += 8) {
      const int r = bj * 32 + tx, c = bi * 32 + k;
      double v = gld(op.src)[r + (long)c * NB];
      if (op.mode == kMoveTransLower && r < c) v = 0.0;
      t[k][tx] = v;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) gst(op.dst)[(bi * 32 + tx) + (long)(bj * 32 + k) * NB] = t[tx][k];
    return;
  }
  for (int k = ty; k < 32; k += 8) {
    const int r = bi * 32 + tx, c = bj * 32 + k;
    const long e = r + (long)c * NB;
    switch (op.mode) {
      case kMoveCopy: gst(op.dst)[e] = gld(op.src)[e]; break;
      case kMoveCopyLower:
        if (r >= c) gst(op.dst)[e] = gld(op.src)[e];
        break;
      case kMoveNegate: gst(op.dst)[e] = -gld(op.src)[e]; break;
      default: gst(op.dst)[e] = (r == c) ? 1.0 : 0.0; break;
    }
  }
}

// *info := 1-based index of the first exact zero on the diagonal of the
// factor (LAPACK dtrtri's singularity check), 0 if none.  One workgroup.
__global__ __launch_bounds__(256) void diag_zero_kernel(int N, int NB, int nt, const double *__restrict__ tiles,
                                                        int *__restrict__ info) {
  __shared__ int best;
  if (threadIdx.x == 0) best = 0x7fffffff;
  __syncthreads();
  int mine = 0x7fffffff;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const int k = i / NB, j = i % NB;
    const double v = tiles[tile_index(k, k, nt) * (long)NB * NB + j + (long)j * NB];
    if (v == 0.0 && i < mine) mine = i;
  }
  atomicMin(&best, mine);
  __syncthreads();
  if (threadIdx.x == 0) *info = (best == 0x7fffffff) ? 0 : best + 1;
}

static hipError_t launch_moves(hipStream_t s, const TileMove *ops, int nops, int NB, const int *d_abort) {
  const int nb = NB / 32;
  for (int o = 0; o < nops; o += 65535) {  // grid y is at most 65535
    hipLaunchKernelGGL(tile_move_kernel, dim3(nb * nb, std::min(65535, nops - o)), dim3(256), 0, s, ops + o, NB,
                       d_abort);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

struct PoinvPlan {
  int N = 0, NB = 0, NT = 0;
  double *tiles = nullptr;
  size_t per = 0;
  int dstride = 0;
  DevBuf d_ops, d_inv, d_q, d_dinv, d_t, d_info;
  // descriptor arrays, all in d_ops at these byte offsets
  size_t o_dinv = 0, o_ident = 0, o_inv = 0, o_trsmr = 0, o_gemm = 0, o_trsml = 0, o_fin = 0, o_trans = 0,
         o_lauum = 0;
  std::vector<int> trsmr_off, trsmr_cnt, gemm_off, gemm_cnt, trsml_off, trsml_cnt, fin_off, fin_cnt;
  int n_trans = 0, n_lauum = 0;
  std::vector<std::pair<int, int>> lauum_mn;  // output tile of each LAUUM op, (-1, -1) for padding
  bool fast = false;

  double *tile(int m, int n) const { return tiles + tile_index(m, n, NT) * (long)per; }

  int build() {
    NT = (N + NB - 1) / NB;
    per = (size_t)NB * NB;
    dstride = ((NB + 31) / 32) * 32 * 32;
    const size_t ntile = (size_t)NT * (NT + 1) / 2;
    TTGC_HIP(d_inv.reserve(sizeof(double) * per * NT));
    TTGC_HIP(d_q.reserve(sizeof(double) * per * NT));
    TTGC_HIP(d_dinv.reserve(sizeof(double) * (size_t)dstride * NT));
    TTGC_HIP(d_info.reserve(sizeof(int)));
    double *inv = static_cast<double *>(d_inv.ptr), *q = static_cast<double *>(d_q.ptr);
    double *dinv = static_cast<double *>(d_dinv.ptr);
    auto S = [&](int k) { return inv + (size_t)k * per; };
    auto Q = [&](int n) { return q + (size_t)n * per; };
    std::vector<PotrfOp> dops;
    std::vector<TileMove> ident, fin, trans;
    std::vector<TrsmOp<double>> invops, trsmr;
    std::vector<UpdOp<double>> gemm, trsml, lauum;
    for (int k = 0; k < NT; ++k) {
      dops.push_back(PotrfOp{tile(k, k), dinv + (size_t)k * dstride, nullptr, NB, NB, 0, 0, nullptr});
      ident.push_back(TileMove{nullptr, S(k), kMoveIdentity});
      invops.push_back(TrsmOp<double>{tile(k, k), S(k), dinv + (size_t)k * dstride, NB, NB, NB, NB});
    }
    for (int k = 0; k < NT; ++k) {
      const double *dk = dinv + (size_t)k * dstride;
      trsmr_off.push_back((int)trsmr.size());
      for (int m = k + 1; m < NT; ++m) trsmr.push_back(TrsmOp<double>{tile(k, k), tile(m, k), dk, NB, NB, NB, NB});
      trsmr_cnt.push_back((int)trsmr.size() - trsmr_off.back());
      gemm_off.push_back((int)gemm.size());
      for (int m = k + 1; m < NT; ++m)
        for (int n = 0; n < k; ++n)
          gemm.push_back(UpdOp<double>{tile(m, k), tile(k, n), tile(m, n), NB, NB, NB, NB, NB, NB, 0});
      gemm_cnt.push_back((int)gemm.size() - gemm_off.back());
      trsml_off.push_back((int)trsml.size());
      for (int n = 0; n < k; ++n) trsml.push_back(UpdOp<double>{S(k), tile(k, n), Q(n), NB, NB, NB, NB, NB, NB, 0});
      trsml_cnt.push_back((int)trsml.size() - trsml_off.back());
      fin_off.push_back((int)fin.size());
      for (int m = k + 1; m < NT; ++m) fin.push_back(TileMove{tile(m, k), tile(m, k), kMoveNegate});
      for (int n = 0; n < k; ++n) fin.push_back(TileMove{Q(n), tile(k, n), kMoveCopy});
      fin.push_back(TileMove{S(k), tile(k, k), kMoveCopyLower});
      fin_cnt.push_back((int)fin.size() - fin_off.back());
    }
    // LAUUM: transposed copy, then one GEMM per output tile (long ones first)
    for (int n = 0; n < NT; ++n)
      for (int m = n; m < NT; ++m) trans.push_back(TileMove{tile(m, n), nullptr, m == n ? kMoveTransLower : kMoveTrans});
    // The update kernel gives XCD x the x-th eighth of the op list, and the
    // depth of op (m, n) is NT - m tiles: deal the ops, deepest first, round
    // robin into 8 equal-length groups (padded with empty ops) so that every
    // XCD gets the same work, each in deepest-first order.
    {
      std::vector<std::pair<int, int>> mn;
      for (int m = 0; m < NT; ++m)
        for (int n = 0; n <= m; ++n) mn.emplace_back(m, n);
      const int len = ((int)mn.size() + kXcds - 1) / kXcds;
      lauum_mn.assign((size_t)len * kXcds, std::make_pair(-1, -1));
      for (size_t i = 0; i < mn.size(); ++i) lauum_mn[(i % kXcds) * len + i / kXcds] = mn[i];
      for (auto &p : lauum_mn)
        lauum.push_back(p.first < 0 ? UpdOp<double>{nullptr, nullptr, tiles, 0, 0, 0, NB, NB, NB, 0}
                                    : UpdOp<double>{nullptr, nullptr, tile(p.first, p.second), NB, NB,
                                                    NB * (NT - p.first), NB, NB, NB, p.first == p.second ? 1 : 0});
    }
    n_trans = (int)trans.size();
    n_lauum = (int)lauum.size();
    (void)ntile;
    fast = update_fast(UpdOp<double>{tiles, tiles, tiles, NB, NB, NB, NB, NB, NB, 0}, NB);
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    size_t off = 0;
    auto place = [&](size_t &o, size_t bytes) {
      o = off;
      off = align(off + bytes);
    };
    place(o_dinv, dops.size() * sizeof(PotrfOp));
    place(o_ident, ident.size() * sizeof(TileMove));
    place(o_inv, invops.size() * sizeof(TrsmOp<double>));
    place(o_trsmr, trsmr.size() * sizeof(TrsmOp<double>));
    place(o_gemm, gemm.size() * sizeof(UpdOp<double>));
    place(o_trsml, trsml.size() * sizeof(UpdOp<double>));
    place(o_fin, fin.size() * sizeof(TileMove));
    place(o_trans, trans.size() * sizeof(TileMove));
    place(o_lauum, lauum.size() * sizeof(UpdOp<double>));
    TTGC_HIP(d_ops.reserve(off));
    std::vector<char> host(off);
    auto put = [&](size_t o, const void *p, size_t bytes) {
      if (bytes) memcpy(host.data() + o, p, bytes);
    };
    put(o_dinv, dops.data(), dops.size() * sizeof(PotrfOp));
    put(o_ident, ident.data(), ident.size() * sizeof(TileMove));
    put(o_inv, invops.data(), invops.size() * sizeof(TrsmOp<double>));
    put(o_trsmr, trsmr.data(), trsmr.size() * sizeof(TrsmOp<double>));
    put(o_gemm, gemm.data(), gemm.size() * sizeof(UpdOp<double>));
    put(o_trsml, trsml.data(), trsml.size() * sizeof(UpdOp<double>));
    put(o_fin, fin.data(), fin.size() * sizeof(TileMove));
    put(o_trans, trans.data(), trans.size() * sizeof(TileMove));
    put(o_lauum, lauum.data(), lauum.size() * sizeof(UpdOp<double>));
    TTGC_HIP(hipMemcpy(d_ops.ptr, host.data(), off, hipMemcpyHostToDevice));
    return TTG_CHOL_OK;
  }

  template <typename Op>
  const Op *at(size_t o, int i = 0) const {
    return reinterpret_cast<const Op *>(static_cast<const char *>(d_ops.ptr) + o) + i;
  }

  // the transposed copy T and its operand pointers (first LAUUM only)
  int ensure_lauum() {
    if (d_t.ptr) return TTG_CHOL_OK;
    const size_t ntile = (size_t)NT * (NT + 1) / 2;
    TTGC_HIP(d_t.reserve(sizeof(double) * per * ntile));
    double *T = static_cast<double *>(d_t.ptr);
    std::vector<TileMove> trans(n_trans);
    std::vector<UpdOp<double>> lauum(n_lauum);
    TTGC_HIP(hipMemcpy(trans.data(), at<TileMove>(o_trans), n_trans * sizeof(TileMove), hipMemcpyDeviceToHost));
    TTGC_HIP(hipMemcpy(lauum.data(), at<UpdOp<double>>(o_lauum), n_lauum * sizeof(UpdOp<double>),
                       hipMemcpyDeviceToHost));
    for (auto &t : trans) t.dst = T + (t.src - tiles);
    for (size_t i = 0; i < lauum_mn.size(); ++i) {
      const int m = lauum_mn[i].first, n = lauum_mn[i].second;
      if (m < 0) continue;
      lauum[i].A = T + tile_index(m, m, NT) * (long)per;
      lauum[i].B = T + tile_index(m, n, NT) * (long)per;
    }
    TTGC_HIP(hipMemcpy(const_cast<TileMove *>(at<TileMove>(o_trans)), trans.data(), n_trans * sizeof(TileMove),
                       hipMemcpyHostToDevice));
    TTGC_HIP(hipMemcpy(const_cast<UpdOp<double> *>(at<UpdOp<double>>(o_lauum)), lauum.data(),
                       n_lauum * sizeof(UpdOp<double>), hipMemcpyHostToDevice));
    return TTG_CHOL_OK;
  }

  int trtri(hipStream_t s, int *info) {
    const UpdOp<double> unone{};
    const TrsmOp<double> tnone{};
    const PotrfOp pnone{};
    TTGC_HIP(launch_prepare(s, N, NB, tiles, info));
    hipLaunchKernelGGL(diag_zero_kernel, dim3(1), dim3(256), 0, s, N, NB, NT, tiles, info);
    TTGC_HIP(hipGetLastError());
    TTGC_HIP(launch_dinv(s, at<PotrfOp>(o_dinv), NT, pnone));
    TTGC_HIP(launch_moves(s, at<TileMove>(o_ident), NT, NB, info));
    TTGC_HIP(launch_trsm<double>(s, at<TrsmOp<double>>(o_inv), NT, tnone, NB, info, false));
    for (int k = 0; k < NT; ++k) {
      TTGC_HIP(launch_trsm<double>(s, at<TrsmOp<double>>(o_trsmr, trsmr_off[k]), trsmr_cnt[k], tnone, NB, info,
                                   false));
      TTGC_HIP(launch_update<double>(s, at<UpdOp<double>>(o_gemm, gemm_off[k]), gemm_cnt[k], unone, NB, NB, -1.0,
                                     1.0, fast, info, true));
      TTGC_HIP(launch_update<double>(s, at<UpdOp<double>>(o_trsml, trsml_off[k]), trsml_cnt[k], unone, NB, NB, 1.0,
                                     0.0, fast, info, true));
      TTGC_HIP(launch_moves(s, at<TileMove>(o_fin, fin_off[k]), fin_cnt[k], NB, info));
    }
    return TTG_CHOL_OK;
  }

  // info: skip everything when nonzero at kernel start (after a failed TRTRI)
  int lauum(hipStream_t s, const int *info) {
    int rc = ensure_lauum();
    if (rc != TTG_CHOL_OK) return rc;
    const UpdOp<double> unone{};
    TTGC_HIP(launch_moves(s, at<TileMove>(o_trans), n_trans, NB, info));
    TTGC_HIP(launch_update<double>(s, at<UpdOp<double>>(o_lauum), n_lauum, unone, NB, NB, 1.0, 0.0, fast, info,
                                   false));
    return TTG_CHOL_OK;
  }

  ~PoinvPlan() {
    for (DevBuf *b : {&d_ops, &d_inv, &d_q, &d_dinv, &d_t, &d_info})
      if (b->ptr) (void)hipFree(b->ptr);
  }
};

static std::mutex g_poinv_mu;
static std::map<std::tuple<int, int, double *>, std::unique_ptr<PoinvPlan>> g_poinv;

void release_poinv_plans() {
  std::lock_guard<std::mutex> lk(g_poinv_mu);
  g_poinv.clear();
}

static int poinv_plan(const char *who, int N, int NB, double *tiles, PoinvPlan **out) {
  if (N < 0 || NB <= 0 || NB > TTG_CHOL_MAX_TILE || (N > 0 && !tiles)) {
    set_error("%s: bad argument (N=%d NB=%d)", who, N, NB);
    return TTG_CHOL_EINVAL;
  }
  if (NB % 32 != 0) {
    set_error("%s: NB=%d must be a multiple of 32", who, NB);
    return TTG_CHOL_ENOTSUP;
  }
  auto key = std::make_tuple(N, NB, tiles);
  auto &plan = g_poinv[key];
  if (!plan) {
    plan.reset(new PoinvPlan);
    plan->N = N;
    plan->NB = NB;
    plan->tiles = tiles;
    int rc = plan->build();
    if (rc != TTG_CHOL_OK) {
      g_poinv.erase(key);
      return rc;
    }
  }
  *out = plan.get();
  return TTG_CHOL_OK;
}

}  // namespace ttgc

using namespace ttgc;

extern "C" int ttg_chol_dtrtri_tiled(void *stream, int N, int NB, double *tiles, int *d_info) {
  clear_error();
  if (N == 0 && NB > 0) return TTG_CHOL_OK;
  std::lock_guard<std::mutex> lk(g_poinv_mu);
  PoinvPlan *p = nullptr;
  int rc = poinv_plan("dtrtri_tiled", N, NB, tiles, &p);
  if (rc != TTG_CHOL_OK) return rc;
  return p->trtri(as_stream(stream), d_info ? d_info : static_cast<int *>(p->d_info.ptr));
}

extern "C" int ttg_chol_dlauum_tiled(void *stream, int N, int NB, double *tiles) {
  clear_error();
  if (N == 0 && NB > 0) return TTG_CHOL_OK;
  std::lock_guard<std::mutex> lk(g_poinv_mu);
  PoinvPlan *p = nullptr;
  int rc = poinv_plan("dlauum_tiled", N, NB, tiles, &p);
  if (rc != TTG_CHOL_OK) return rc;
  hipStream_t s = as_stream(stream);
  TTGC_HIP(launch_prepare(s, N, NB, tiles, static_cast<int *>(p->d_info.ptr)));
  return p->lauum(s, nullptr);
}

extern "C" int ttg_chol_dpotri_tiled(void *stream, int N, int NB, double *tiles, int *d_info) {
  clear_error();
  if (N == 0 && NB > 0) return TTG_CHOL_OK;
  std::lock_guard<std::mutex> lk(g_poinv_mu);
  PoinvPlan *p = nullptr;
  int rc = poinv_plan("dpotri_tiled", N, NB, tiles, &p);
  if (rc != TTG_CHOL_OK) return rc;
  hipStream_t s = as_stream(stream);
  int *info = d_info ? d_info : static_cast<int *>(p->d_info.ptr);
  rc = p->trtri(s, info);
  if (rc != TTG_CHOL_OK) return rc;
  return p->lauum(s, info);
}
