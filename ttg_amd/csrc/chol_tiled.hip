// Tiled FP64 Cholesky driver: the task graph of examples/potrf/potrf.h with
// the launches grouped per panel step.
//
// Step k of the graph consists of POTRF(k), TRSM(m,k) for m > k, and the
// updates SYRK(k,m) / GEMM(m,n,k) for k < n <= m.  The plan fuses every
// TRSM of a step into one launch and every SYRK+GEMM of a step into (at most)
// two launches:
//   updA(k): the tiles of column k+1, which feed POTRF/TRSM(k+1);
//   updB(k): the rest of the trailing matrix.
// With lookahead, POTRF(k+1) and TRSM(k+1) run on a second stream as soon as
// updA(k) is done, overlapping updB(k).  Per tile the update order over k is
// the one the reference graph fixes (GEMM(m,n,0), GEMM(m,n,1), ...), so the
// arithmetic per tile is the same as the reference's.  The whole launch
// sequence is captured once into a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#include "chol_internal.hpp"
#include "chol_kernels.hpp"
#include "chol_profile.hpp"
#include "chol_tiled.hpp"

namespace ttgc {

struct TiledPlan {
  int N = 0, NB = 0, NT = 0;
  double *tiles = nullptr;
  unsigned flags = 0;
  DevBuf d_ops, d_dinv, d_info;
  size_t o_potrf = 0, o_trsm = 0, o_upd = 0;
  std::vector<int> trsm_off, trsm_cnt, updA_off, updA_cnt, updB_off, updB_cnt;
  hipStream_t s_main = nullptr, s_look = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr, ev_panel = nullptr, ev_col = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int *bound_info = nullptr;  // the d_info the captured graph writes
  int *desc_info = nullptr;   // the d_info stored in the POTRF descriptors

  ~TiledPlan() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t e : {ev_in, ev_out, ev_panel, ev_col})
      if (e) (void)hipEventDestroy(e);
    if (s_main) (void)hipStreamDestroy(s_main);
    if (s_look) (void)hipStreamDestroy(s_look);
    for (DevBuf *b : {&d_ops, &d_dinv, &d_info})
      if (b->ptr) (void)hipFree(b->ptr);
  }

  PotrfOp *potrf_ops() { return reinterpret_cast<PotrfOp *>(static_cast<char *>(d_ops.ptr) + o_potrf); }
  TrsmOp<double> *trsm_ops() {
    return reinterpret_cast<TrsmOp<double> *>(static_cast<char *>(d_ops.ptr) + o_trsm);
  }
  UpdOp<double> *upd_ops() { return reinterpret_cast<UpdOp<double> *>(static_cast<char *>(d_ops.ptr) + o_upd); }

  double *tile(int m, int n) { return tiles + tile_index(m, n, NT) * (long)NB * NB; }

  int build(int *info_ptr) {
    NT = (N + NB - 1) / NB;
    const int dstride = ((NB + 31) / 32) * 32 * 32;
    TTGC_HIP(d_dinv.reserve(sizeof(double) * dstride));
    double *dinv = static_cast<double *>(d_dinv.ptr);
    std::vector<PotrfOp> potrf;
    std::vector<TrsmOp<double>> trsm;
    std::vector<UpdOp<double>> upd;
    // every tile is a full NB x NB block once the edge padding is set
    for (int k = 0; k < NT; ++k) {
      potrf.push_back(PotrfOp{tile(k, k), dinv, info_ptr, NB, NB, k * NB, 0});
      trsm_off.push_back((int)trsm.size());
      for (int m = k + 1; m < NT; ++m) trsm.push_back(TrsmOp<double>{tile(k, k), tile(m, k), dinv, NB, NB, NB, NB});
      trsm_cnt.push_back((int)trsm.size() - trsm_off.back());
      // updA(k): column k+1 -- SYRK(k,k+1) then GEMM(m,k+1,k)
      updA_off.push_back((int)upd.size());
      if (k + 1 < NT)
        for (int m = k + 1; m < NT; ++m)
          upd.push_back(UpdOp<double>{tile(m, k), tile(k + 1, k), tile(m, k + 1), NB, NB, NB, NB, NB, NB,
                                      m == k + 1 ? 1 : 0});
      updA_cnt.push_back((int)upd.size() - updA_off.back());
      // updB(k): columns k+2..NT-1
      updB_off.push_back((int)upd.size());
      for (int n = k + 2; n < NT; ++n)
        for (int m = n; m < NT; ++m)
          upd.push_back(UpdOp<double>{tile(m, k), tile(n, k), tile(m, n), NB, NB, NB, NB, NB, NB, m == n ? 1 : 0});
      updB_cnt.push_back((int)upd.size() - updB_off.back());
    }
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    o_potrf = 0;
    o_trsm = align(potrf.size() * sizeof(PotrfOp));
    o_upd = align(o_trsm + trsm.size() * sizeof(TrsmOp<double>));
    const size_t total = align(o_upd + upd.size() * sizeof(UpdOp<double>));
    TTGC_HIP(d_ops.reserve(total));
    std::vector<char> host(total);
    memcpy(host.data() + o_potrf, potrf.data(), potrf.size() * sizeof(PotrfOp));
    if (!trsm.empty()) memcpy(host.data() + o_trsm, trsm.data(), trsm.size() * sizeof(TrsmOp<double>));
    if (!upd.empty()) memcpy(host.data() + o_upd, upd.data(), upd.size() * sizeof(UpdOp<double>));
    TTGC_HIP(hipMemcpy(d_ops.ptr, host.data(), total, hipMemcpyHostToDevice));
    TTGC_HIP(hipStreamCreateWithFlags(&s_main, hipStreamNonBlocking));
    TTGC_HIP(hipStreamCreateWithFlags(&s_look, hipStreamNonBlocking));
    for (hipEvent_t *e : {&ev_in, &ev_out, &ev_panel, &ev_col})
      TTGC_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    return TTG_CHOL_OK;
  }

  // Enqueue the whole factorization on s_main (and s_look), writing *info.
  // With `prof`, every launch is bracketed by timing events.
  int enqueue(int *info, bool prof) {
    const bool fast = NB % 128 == 0;
    const double m1 = -1.0, p1 = 1.0;
    const double b3 = (double)NB * NB * NB;
    UpdOp<double> unone{};
    TrsmOp<double> tnone{};
    PotrfOp pnone{};
    TTGC_HIP(launch_prepare(s_main, N, NB, tiles, info));
    const bool la = (flags & TTG_CHOL_F_LOOKAHEAD) != 0;
    hipStream_t sp = la ? s_look : s_main;
    if (la) TTGC_HIP(hipEventRecord(ev_col, s_main));
    auto upd = [&](int off, int cnt, int nsyrk) -> int {
      const int slot = prof ? prof_begin(s_main) : -1;
      TTGC_HIP(launch_update<double>(s_main, upd_ops() + off, cnt, unone, NB, NB, m1, p1, fast, info));
      if (prof) prof_end(slot, s_main, PROF_UPDATE, (cnt - nsyrk) * 2.0 * b3 + nsyrk * (double)NB * NB * (NB + 1));
      return TTG_CHOL_OK;
    };
    for (int k = 0; k < NT; ++k) {
      if (la) TTGC_HIP(hipStreamWaitEvent(s_look, ev_col, 0));
      int slot = prof ? prof_begin(sp) : -1;
      TTGC_HIP(launch_potrf(sp, potrf_ops() + k, 1, pnone, info));
      if (prof) prof_end(slot, sp, PROF_POTRF, b3 / 3.0);
      if (trsm_cnt[k]) {
        slot = prof ? prof_begin(sp) : -1;
        TTGC_HIP(launch_trsm<double>(sp, trsm_ops() + trsm_off[k], trsm_cnt[k], tnone, NB, info));
        if (prof) prof_end(slot, sp, PROF_TRSM, trsm_cnt[k] * b3);
      }
      if (la) {
        TTGC_HIP(hipEventRecord(ev_panel, s_look));
        TTGC_HIP(hipStreamWaitEvent(s_main, ev_panel, 0));
      }
      if (updA_cnt[k]) {
        int rc = upd(updA_off[k], updA_cnt[k], 1);
        if (rc) return rc;
      }
      if (la) TTGC_HIP(hipEventRecord(ev_col, s_main));
      if (updB_cnt[k]) {
        int rc = upd(updB_off[k], updB_cnt[k], NT - k - 2);
        if (rc) return rc;
      }
    }
    if (la) {
      TTGC_HIP(hipEventRecord(ev_panel, s_look));
      TTGC_HIP(hipStreamWaitEvent(s_main, ev_panel, 0));
    }
    return TTG_CHOL_OK;
  }

  int run(hipStream_t user, int *info) {
    const bool prof = (flags & TTG_CHOL_F_PROFILE) != 0;
    if ((flags & TTG_CHOL_F_GRAPH) && !prof) {
      if (!gexec || bound_info != info) {
        if (gexec) {
          TTGC_HIP(hipGraphExecDestroy(gexec));
          gexec = nullptr;
        }
        if (graph) {
          TTGC_HIP(hipGraphDestroy(graph));
          graph = nullptr;
        }
        TTGC_HIP(hipStreamBeginCapture(s_main, hipStreamCaptureModeThreadLocal));
        int rc = enqueue(info, false);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s_main, &g);
        if (rc != TTG_CHOL_OK) {
          if (g) (void)hipGraphDestroy(g);
          return rc;
        }
        TTGC_HIP(e);
        graph = g;
        TTGC_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        bound_info = info;
      }
      TTGC_HIP(hipGraphLaunch(gexec, user));
      return TTG_CHOL_OK;
    }
    TTGC_HIP(hipEventRecord(ev_in, user));
    TTGC_HIP(hipStreamWaitEvent(s_main, ev_in, 0));
    int rc = enqueue(info, prof);
    if (rc != TTG_CHOL_OK) return rc;
    TTGC_HIP(hipEventRecord(ev_out, s_main));
    TTGC_HIP(hipStreamWaitEvent(user, ev_out, 0));
    return TTG_CHOL_OK;
  }
};

static std::mutex g_plan_mu;
static std::map<std::tuple<int, int, double *, unsigned>, std::unique_ptr<TiledPlan>> g_plans;

}  // namespace ttgc

using namespace ttgc;

extern "C" {

int ttg_chol_dpotrf_tiled(void *stream, int N, int NB, double *tiles, int *d_info, unsigned flags) {
  clear_error();
  if (N < 0 || NB <= 0 || NB > TTG_CHOL_MAX_TILE || (N > 0 && !tiles)) {
    set_error("dpotrf_tiled: bad argument (N=%d NB=%d)", N, NB);
    return TTG_CHOL_EINVAL;
  }
  if (NB % 32 != 0) {
    set_error("dpotrf_tiled: NB=%d must be a multiple of 32", NB);
    return TTG_CHOL_ENOTSUP;
  }
  if (N == 0) return TTG_CHOL_OK;
  std::lock_guard<std::mutex> lk(g_plan_mu);
  auto key = std::make_tuple(N, NB, tiles, flags);
  auto &plan = g_plans[key];
  if (!plan) {
    plan.reset(new TiledPlan);
    plan->N = N;
    plan->NB = NB;
    plan->tiles = tiles;
    plan->flags = flags;
    TTGC_HIP(plan->d_info.reserve(sizeof(int)));
    int rc = plan->build(nullptr);
    if (rc != TTG_CHOL_OK) {
      g_plans.erase(key);
      return rc;
    }
  }
  int *info = d_info ? d_info : static_cast<int *>(plan->d_info.ptr);
  // the POTRF descriptors carry the info pointer: rewrite them if it changed
  if (plan->desc_info != info) {
    std::vector<PotrfOp> ops(plan->NT);
    TTGC_HIP(hipMemcpy(ops.data(), plan->potrf_ops(), sizeof(PotrfOp) * plan->NT, hipMemcpyDeviceToHost));
    for (auto &o : ops) o.info = info;
    TTGC_HIP(hipMemcpy(plan->potrf_ops(), ops.data(), sizeof(PotrfOp) * plan->NT, hipMemcpyHostToDevice));
    plan->desc_info = info;
  }
  return plan->run(as_stream(stream), info);
}

}  // extern "C"
