// Plan of the tiled factorization (the examples/potrf task graph with the
// launches grouped per panel step), templated on the storage type T of the
// factor tiles: double (plain FP64) or float (mixed precision: FP32 TRSM and
// trailing updates, or FP16 / BF16 updates; the diagonal tiles in FP64, or in
// FP32 under the one-product FP16 / BF16 updates).
//
// Step k of the graph is POTRF(k), TRSM(m,k) for m > k, and the updates
// SYRK(k,m) / GEMM(m,n,k) for k < n <= m.  Every TRSM of a step is one
// launch; the updates of a step are two launches:
//   updA(k): the tiles of column k+1, which feed POTRF/TRSM(k+1);
//   updB(k): the rest of the trailing matrix.
// With lookahead, POTRF(k+1) and TRSM(k+1) run on a second stream as soon as
// updA(k) is done, overlapping updB(k).  Per tile the update order over k is
// the one the reference graph fixes (GEMM(m,n,0), GEMM(m,n,1), ...).  The
// launch sequence is captured once into a hipGraph and replayed.
//
// Reserved CUs (TTG_CHOL_POTRF_CUS; default 8 for float tiles, 0 for FP64,
// whose long update launches measured slower on 248 CUs than the serialized
// POTRF costs): the single-workgroup POTRF
// gets a stream of its own whose queue may only use the first R CUs of the
// device mask, and the update stream the others.  KFD interleaves the mask
// bits over the XCDs (bit b -> XCD b % 8), so R = 8 is one CU per XCD.
// POTRF(k+1) then runs beside updB(k) at full speed instead of after it.
// The masks live on the HSA queues, which a replayed hipGraph does not use,
// so this schedule is launched directly, without graph capture.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

#include "chol_f16.hpp"
#include "chol_internal.hpp"
#include "chol_kernels.hpp"
#include "chol_profile.hpp"
#include "chol_tiled.hpp"

namespace ttgc {

template <typename T>
hipError_t launch_prepare_t(hipStream_t s, int N, int NB, T *tiles, int *d_info);

template <typename T>
struct TiledPlanT {
  int N = 0, NB = 0, NT = 0;
  T *tiles = nullptr;
  unsigned flags = 0;
  int lowp = 0;  // float tiles only: 0 FP32 updates, else kLowF16x2 / kLowF16 / kLowBF16
  // The diagonal tiles are factored in FP64, except under the one-product
  // FP16 / BF16 updates, whose operand rounding (2^-11 / 2^-8 of each row's
  // largest entry) dwarfs an FP32 factorization's: there POTRF runs in FP32
  // on the tile itself (FP32 MFMA, half the bytes of its left-looking reads).
  bool potrf_in_ts() const { return lowp == kLowF16 || lowp == kLowBF16; }
  DevBuf d_ops, d_dinv, d_info, d_pw;
  DevBuf d_panel, d_scale, d_f16ops;  // split-FP16 panels (two slots, 2G when grouped), scales, descriptors
  std::vector<int> split_off, split_cnt;
  // grouped schedule (split-FP16 modes with lookahead): the steps in groups
  // of G; per step k of a group, the eager update of the group's later
  // columns from panel k; per group, the G-panel updates of the next group's
  // columns (bulkA) and of all further columns (bulkB)
  bool paired = false;
  int G = 2;
  // Optional work the caller's first step leaves to the plan: enqueued on the
  // bulk stream before anything there touches the tiles of columns >=
  // first_cols() (the grouped schedule: after the panel chain of the first
  // group has been released, so that the two overlap; otherwise at the
  // start).  Set by the caller for one run.
  std::function<int(hipStream_t)> pre_bulk;
  // columns the caller must have ready before run(): the first group's in
  // the grouped schedule with full tiles, else all of them
  // (the graph-replayed schedule, used without reserved CUs, takes all)
  int first_cols() const { return (paired && N % NB == 0 && s_diag) ? std::min(G, NT) : NT; }
  int TG = 2;  // side of the square tile groups of the bulk updates (TTG_CHOL_TILE_GROUP)
  std::vector<int> pe_off, pe_cnt, pa_off, pa_cnt, pb_off, pb_cnt;  // eager per step; bulkA, bulkB per group
  hipEvent_t ev_bulk = nullptr;
  size_t o_split = 0, o_f16 = 0;
  size_t o_potrf = 0, o_trsm = 0, o_upd = 0;
  std::vector<int> trsm_off, trsm_cnt, updA_off, updA_cnt, updB_off, updB_cnt;
  hipStream_t s_main = nullptr, s_look = nullptr, s_diag = nullptr;  // s_diag: POTRF on reserved CUs
  hipEvent_t ev_in = nullptr, ev_out = nullptr, ev_panel = nullptr, ev_col = nullptr, ev_diag = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int *bound_info = nullptr;  // the d_info the captured graph writes
  int *desc_info = nullptr;   // the d_info stored in the POTRF descriptors

  ~TiledPlanT() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t e : {ev_in, ev_out, ev_panel, ev_col, ev_diag, ev_bulk})
      if (e) (void)hipEventDestroy(e);
    if (s_main) (void)hipStreamDestroy(s_main);
    if (s_look) (void)hipStreamDestroy(s_look);
    if (s_diag) (void)hipStreamDestroy(s_diag);
    for (DevBuf *b : {&d_ops, &d_dinv, &d_info, &d_pw, &d_panel, &d_scale, &d_f16ops})
      if (b->ptr) (void)hipFree(b->ptr);
  }

  PotrfOpT<T> *potrf_ops() { return reinterpret_cast<PotrfOpT<T> *>(static_cast<char *>(d_ops.ptr) + o_potrf); }
  TrsmOp<T> *trsm_ops() { return reinterpret_cast<TrsmOp<T> *>(static_cast<char *>(d_ops.ptr) + o_trsm); }
  UpdOp<T> *upd_ops() { return reinterpret_cast<UpdOp<T> *>(static_cast<char *>(d_ops.ptr) + o_upd); }
  T *tile(int m, int n) { return tiles + tile_index(m, n, NT) * (long)NB * NB; }

  // Tiles (m, n), k + 2 <= n <= m < NT, of updB(k) in launch order: TG x TG
  // groups of tiles (default 2), so that the workgroups running together on
  // one XCD share TG row panels and TG column panels (measured 6% less time
  // per update launch than column-by-column order with an earlier update
  // kernel; with the current one, sides 1 to 4 give the same solve time
  // within noise).  The order within a launch does not change any tile's
  // update order over k.
  template <typename F>
  void updB_tiles(int k, F &&f) const {
    for (int n0 = k + 2; n0 < NT; n0 += TG)
      for (int m0 = n0; m0 < NT; m0 += TG)
        for (int dn = 0; dn < TG; ++dn)
          for (int dm = 0; dm < TG; ++dm) {
            const int m = m0 + dm, n = n0 + dn;
            if (m < NT && n < NT && m >= n) f(m, n);
          }
  }

  int build() {
    NT = (N + NB - 1) / NB;
    TG = std::max(1, std::min(8, plan_knob(TTG_CHOL_KNOB_TILE_GROUP, "TTG_CHOL_TILE_GROUP", TG)));
    const int dstride = ((NB + 31) / 32) * 32 * 32;
    TTGC_HIP(d_dinv.reserve(sizeof(T) * dstride));
    TTGC_HIP(d_info.reserve(sizeof(int)));
    TTGC_HIP(d_pw.reserve(sizeof(double) * (size_t)NB * NB));
    T *dinv = static_cast<T *>(d_dinv.ptr);
    double *pw = sizeof(T) == 8 ? nullptr : static_cast<double *>(d_pw.ptr);
    std::vector<PotrfOpT<T>> potrf;
    std::vector<TrsmOp<T>> trsm;
    std::vector<UpdOp<T>> upd;
    // every tile is a full NB x NB block once the edge padding is set
    for (int k = 0; k < NT; ++k) {
      potrf.push_back(PotrfOpT<T>{tile(k, k), dinv, nullptr, NB, NB, k * NB, 0, pw});
      trsm_off.push_back((int)trsm.size());
      for (int m = k + 1; m < NT; ++m) trsm.push_back(TrsmOp<T>{tile(k, k), tile(m, k), dinv, NB, NB, NB, NB});
      trsm_cnt.push_back((int)trsm.size() - trsm_off.back());
      updA_off.push_back((int)upd.size());
      if (k + 1 < NT)
        for (int m = k + 1; m < NT; ++m)
          upd.push_back(UpdOp<T>{tile(m, k), tile(k + 1, k), tile(m, k + 1), NB, NB, NB, NB, NB, NB, m == k + 1 ? 1 : 0});
      updA_cnt.push_back((int)upd.size() - updA_off.back());
      updB_off.push_back((int)upd.size());
      updB_tiles(k, [&](int m, int n) {
        upd.push_back(UpdOp<T>{tile(m, k), tile(n, k), tile(m, n), NB, NB, NB, NB, NB, NB, m == n ? 1 : 0});
      });
      updB_cnt.push_back((int)upd.size() - updB_off.back());
    }
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    o_potrf = 0;
    o_trsm = align(potrf.size() * sizeof(PotrfOpT<T>));
    o_upd = align(o_trsm + trsm.size() * sizeof(TrsmOp<T>));
    const size_t total = align(o_upd + upd.size() * sizeof(UpdOp<T>));
    TTGC_HIP(d_ops.reserve(total));
    std::vector<char> host(total);
    memcpy(host.data() + o_potrf, potrf.data(), potrf.size() * sizeof(PotrfOpT<T>));
    if (!trsm.empty()) memcpy(host.data() + o_trsm, trsm.data(), trsm.size() * sizeof(TrsmOp<T>));
    if (!upd.empty()) memcpy(host.data() + o_upd, upd.data(), upd.size() * sizeof(UpdOp<T>));
    TTGC_HIP(hipMemcpy(d_ops.ptr, host.data(), total, hipMemcpyHostToDevice));
    if (lowp) {
      // groups of 3 steps by default: with the 8-wave POTRF the panel chain
      // is short enough that a third panel per bulk update pays (measured
      // 431 -> 425 ms per solve at N=65536; groups of 4 were no faster)
      G = std::max(2, std::min(8, plan_knob(TTG_CHOL_KNOB_GROUP, "TTG_CHOL_GROUP", 3)));
      paired = (flags & TTG_CHOL_F_LOOKAHEAD) && NT > G && plan_knob(TTG_CHOL_KNOB_PAIRED, "TTG_CHOL_PAIRED", 1) != 0;
      int rc = paired ? build_f16_paired() : build_f16(upd);
      if (rc) return rc;
    }
    if (flags & TTG_CHOL_F_LOOKAHEAD) reserve_cus();
    if (!s_main) TTGC_HIP(hipStreamCreateWithFlags(&s_main, hipStreamNonBlocking));
    TTGC_HIP(hipStreamCreateWithFlags(&s_look, hipStreamNonBlocking));
    for (hipEvent_t *e : {&ev_in, &ev_out, &ev_panel, &ev_col, &ev_diag, &ev_bulk})
      TTGC_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    return TTG_CHOL_OK;
  }

  // POTRF stream on CUs 0 .. R-1 of the mask, update stream on the rest; on
  // any failure both stay unset and the plain two-stream schedule is used.
  void reserve_cus() {
    const int want = plan_knob(TTG_CHOL_KNOB_POTRF_CUS, "TTG_CHOL_POTRF_CUS", sizeof(T) == 4 ? 8 : 0);
    hipStream_t d = nullptr, m = nullptr;
    if (make_split_streams(want, &d, &m)) {
      s_diag = d;
      s_main = m;
    }
  }

  // Split-FP16 descriptors: panel k lives in slot k % 2 (tile m at position
  // m - k - 1); the update descriptors mirror the FP32 ones one to one.
  SplitOp *split_ops() { return reinterpret_cast<SplitOp *>(static_cast<char *>(d_f16ops.ptr) + o_split); }
  F16UpdOp *f16_ops() { return reinterpret_cast<F16UpdOp *>(static_cast<char *>(d_f16ops.ptr) + o_f16); }

  int build_f16(const std::vector<UpdOp<T>> &upd) {
    const size_t per = (size_t)NB * NB;  // halves per part
    const size_t slot_tiles = NT > 1 ? NT - 1 : 1;
    TTGC_HIP(d_panel.reserve(2 * slot_tiles * 2 * per * sizeof(uint16_t)));
    TTGC_HIP(d_scale.reserve(2 * slot_tiles * NB * sizeof(float)));
    uint16_t *panel = static_cast<uint16_t *>(d_panel.ptr);
    float *scale = static_cast<float *>(d_scale.ptr);
    auto hi = [&](int k, int m) { return panel + ((size_t)(k % 2) * slot_tiles + (m - k - 1)) * 2 * per; };
    auto sc = [&](int k, int m) { return scale + ((size_t)(k % 2) * slot_tiles + (m - k - 1)) * NB; };
    std::vector<SplitOp> sp;
    std::vector<F16UpdOp> fu;
    for (int k = 0; k < NT; ++k) {
      split_off.push_back((int)sp.size());
      for (int m = k + 1; m < NT; ++m)
        sp.push_back(SplitOp{reinterpret_cast<const float *>(tile(m, k)), hi(k, m), hi(k, m) + per, sc(k, m)});
      split_cnt.push_back((int)sp.size() - split_off.back());
    }
    // same order as the FP32 update descriptors (updA(k) then updB(k))
    for (int k = 0; k < NT; ++k) {
      auto add = [&](int m, int n) {
        fu.push_back(F16UpdOp{hi(k, m), hi(k, m) + per, hi(k, n), hi(k, n) + per, sc(k, m), sc(k, n),
                              reinterpret_cast<float *>(tile(m, n)), NB, NB, NB, NB, m == n ? 1 : 0});
      };
      if (k + 1 < NT)
        for (int m = k + 1; m < NT; ++m) add(m, k + 1);
      updB_tiles(k, add);
    }
    if (fu.size() != upd.size()) {
      set_error("internal: split-FP16 descriptor count mismatch");
      return TTG_CHOL_ESTATE;
    }
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    o_split = 0;
    o_f16 = align(sp.size() * sizeof(SplitOp));
    const size_t total = align(o_f16 + fu.size() * sizeof(F16UpdOp));
    TTGC_HIP(d_f16ops.reserve(total));
    std::vector<char> host(total);
    if (!sp.empty()) memcpy(host.data() + o_split, sp.data(), sp.size() * sizeof(SplitOp));
    if (!fu.empty()) memcpy(host.data() + o_f16, fu.data(), fu.size() * sizeof(F16UpdOp));
    TTGC_HIP(hipMemcpy(d_f16ops.ptr, host.data(), total, hipMemcpyHostToDevice));
    return TTG_CHOL_OK;
  }

  // Split-FP16 descriptors of the grouped schedule: panel k in slot k % 2G, so
  // a group's G panels fill G consecutive slots (groups start at multiples
  // of G) and stay intact while the next group splits into the other G.
  // Panel k0 + j of a group then sits a fixed stride after panel k0 for
  // every tile (the slot advances by one, the tile position by minus one),
  // which is what a G-panel update descriptor carries.  Its products are
  // subtracted in panel order, so every tile still takes its updates in the
  // order k = 0, 1, ...
  int build_f16_paired() {
    const size_t per = (size_t)NB * NB;
    const size_t slot_tiles = NT > 1 ? NT - 1 : 1;
    const int S = 2 * G;
    TTGC_HIP(d_panel.reserve(S * slot_tiles * 2 * per * sizeof(uint16_t)));
    TTGC_HIP(d_scale.reserve(S * slot_tiles * NB * sizeof(float)));
    uint16_t *panel = static_cast<uint16_t *>(d_panel.ptr);
    float *scale = static_cast<float *>(d_scale.ptr);
    auto hi = [&](int k, int m) { return panel + ((size_t)(k % S) * slot_tiles + (m - k - 1)) * 2 * per; };
    auto sc = [&](int k, int m) { return scale + ((size_t)(k % S) * slot_tiles + (m - k - 1)) * NB; };
    const long pstride = (long)(slot_tiles - 1) * 2 * (long)per, sstride = (long)(slot_tiles - 1) * NB;
    std::vector<SplitOp> sp;
    std::vector<F16UpdOp> fu;
    for (int k = 0; k < NT; ++k) {
      split_off.push_back((int)sp.size());
      for (int m = k + 1; m < NT; ++m)
        sp.push_back(SplitOp{reinterpret_cast<const float *>(tile(m, k)), hi(k, m), hi(k, m) + per, sc(k, m)});
      split_cnt.push_back((int)sp.size() - split_off.back());
    }
    // tile (m, n) updated by panels k .. k + np - 1
    auto op = [&](int k, int np, int m, int n) {
      fu.push_back(F16UpdOp{hi(k, m), hi(k, m) + per, hi(k, n), hi(k, n) + per, sc(k, m), sc(k, n),
                            reinterpret_cast<float *>(tile(m, n)), NB, NB, NB, NB, m == n ? 1 : 0, np,
                            np > 1 ? pstride : 0, np > 1 ? sstride : 0});
    };
    for (int k0 = 0; k0 < NT; k0 += G) {
      const int ge = std::min(k0 + G, NT);  // the group's steps: k0 .. ge - 1
      for (int k = k0; k < ge; ++k) {
        pe_off.push_back((int)fu.size());
        for (int n = k + 1; n < ge; ++n)
          for (int m = n; m < NT; ++m) op(k, 1, m, n);
        pe_cnt.push_back((int)fu.size() - pe_off.back());
      }
      const int np = ge - k0;
      pa_off.push_back((int)fu.size());
      for (int n = ge; n < std::min(ge + G, NT); ++n)
        for (int m = n; m < NT; ++m) op(k0, np, m, n);
      pa_cnt.push_back((int)fu.size() - pa_off.back());
      pb_off.push_back((int)fu.size());
      if (ge + G < NT) updB_tiles(ge + G - 2, [&](int m, int n) { op(k0, np, m, n); });  // columns >= ge + G
      pb_cnt.push_back((int)fu.size() - pb_off.back());
    }
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    o_split = 0;
    o_f16 = align(sp.size() * sizeof(SplitOp));
    const size_t total = align(o_f16 + fu.size() * sizeof(F16UpdOp));
    TTGC_HIP(d_f16ops.reserve(total));
    std::vector<char> host(total);
    if (!sp.empty()) memcpy(host.data() + o_split, sp.data(), sp.size() * sizeof(SplitOp));
    if (!fu.empty()) memcpy(host.data() + o_f16, fu.data(), fu.size() * sizeof(F16UpdOp));
    TTGC_HIP(hipMemcpy(d_f16ops.ptr, host.data(), total, hipMemcpyHostToDevice));
    return TTG_CHOL_OK;
  }

  // The grouped schedule (stream by stream), group g = steps k0 .. k0+G-1:
  //   s_diag: POTRF(k0) after bulkA(g-1); POTRF(k) after the eager update of step k-1
  //   s_look: per step, TRSM + split, then the eager update of the group's later columns
  //   s_main: bulkA(g) after the group's panels, then bulkB(g)
  // bulkB(g-1) overlaps the whole panel chain of group g; it writes columns
  // >= k0 + G and reads the panels of group g-1, which the chain of group g
  // neither writes nor splits into.
  int enqueue_paired(int *info, bool prof, bool hook) {
    const double b3 = (double)NB * NB * NB;
    PotrfOpT<T> pnone{};
    TrsmOp<T> tnone{};
    TTGC_HIP(launch_prepare_t<T>(s_main, N, NB, tiles, info));
    hipStream_t sd = s_diag ? s_diag : s_look;
    auto dep = [&](hipEvent_t e, hipStream_t from, hipStream_t to) -> int {
      if (from == to) return TTG_CHOL_OK;
      TTGC_HIP(hipEventRecord(e, from));
      TTGC_HIP(hipStreamWaitEvent(to, e, 0));
      return TTG_CHOL_OK;
    };
    auto potrf = [&](int k) -> int {
      const int slot = prof ? prof_begin(sd) : -1;
      TTGC_HIP(launch_potrf<T>(sd, potrf_ops() + k, 1, pnone, info, potrf_in_ts()));
      if (prof) prof_end(slot, sd, PROF_POTRF, b3 / 3.0);
      return dep(ev_diag, sd, s_look);
    };
    auto panel = [&](int k) -> int {
      if (!trsm_cnt[k]) return TTG_CHOL_OK;
      const int slot = prof ? prof_begin(s_look) : -1;
      TTGC_HIP(launch_trsm<T>(s_look, trsm_ops() + trsm_off[k], trsm_cnt[k], tnone, NB, info));
      if (prof) prof_end(slot, s_look, PROF_TRSM, trsm_cnt[k] * b3);
      TTGC_HIP(launch_split(s_look, split_ops() + split_off[k], split_cnt[k], NB, lowp));
      return TTG_CHOL_OK;
    };
    // tile products of a launch: GEMM tiles 2 b^3, SYRK tiles b^2 (b+1), per panel
    auto upd = [&](hipStream_t st, int off, int cnt, int nsyrk, int npan) -> int {
      if (!cnt) return TTG_CHOL_OK;
      const int slot = prof ? prof_begin(st) : -1;
      TTGC_HIP(launch_update_f16x2(st, f16_ops() + off, cnt, NB, info, lowp));
      if (prof)
        prof_end(slot, st, PROF_UPDATE, npan * ((cnt - nsyrk) * 2.0 * b3 + nsyrk * (double)NB * NB * (NB + 1)));
      return TTG_CHOL_OK;
    };
    if (int rc = dep(ev_col, s_main, sd)) return rc;
    if (hook && pre_bulk)
      if (int rc = pre_bulk(s_main)) return rc;
    for (int g = 0, k0 = 0; k0 < NT; ++g, k0 += G) {
      const int ge = std::min(k0 + G, NT);
      if (g > 0) TTGC_HIP(hipStreamWaitEvent(sd, ev_bulk, 0));  // the group's columns complete (bulkA(g-1))
      for (int k = k0; k < ge; ++k) {
        if (k > k0)
          if (int rc = dep(ev_col, s_look, sd)) return rc;  // column k complete (eager update of step k-1)
        if (int rc = potrf(k)) return rc;
        if (int rc = panel(k)) return rc;
        if (int rc = upd(s_look, pe_off[k], pe_cnt[k], ge - k - 1, 1)) return rc;
      }
      if (ge >= NT) break;
      if (int rc = dep(ev_panel, s_look, s_main)) return rc;
      if (int rc = upd(s_main, pa_off[g], pa_cnt[g], std::min(G, NT - ge), ge - k0)) return rc;
      TTGC_HIP(hipEventRecord(ev_bulk, s_main));
      const int nb = NT - ge - G;  // columns of bulkB: its SYRK tiles
      if (int rc = upd(s_main, pb_off[g], pb_cnt[g], nb > 0 ? nb : 0, ge - k0)) return rc;
    }
    if (int rc = dep(ev_panel, s_look, s_main)) return rc;
    if (int rc = dep(ev_diag, sd, s_main)) return rc;
    return TTG_CHOL_OK;
  }

  // Point the POTRF descriptors at `info` (host round trip only when it changes).
  int bind_info(int *info) {
    if (desc_info == info) return TTG_CHOL_OK;
    std::vector<PotrfOpT<T>> ops(NT);
    TTGC_HIP(hipMemcpy(ops.data(), potrf_ops(), sizeof(PotrfOpT<T>) * NT, hipMemcpyDeviceToHost));
    for (auto &o : ops) o.info = info;
    TTGC_HIP(hipMemcpy(potrf_ops(), ops.data(), sizeof(PotrfOpT<T>) * NT, hipMemcpyHostToDevice));
    desc_info = info;
    return TTG_CHOL_OK;
  }

  // Enqueue the whole factorization on s_main (and s_look), writing *info.
  // With `prof`, every launch is bracketed by timing events.
  int enqueue(int *info, bool prof) {
    const bool late = paired && N % NB == 0;  // the grouped schedule runs the hook after the chain's release
    if (pre_bulk && !late)
      if (int rc = pre_bulk(s_main)) return rc;
    if (paired) return enqueue_paired(info, prof, late);
    const bool fast = NB % 128 == 0;
    const T m1 = T(-1), p1 = T(1);
    const double b3 = (double)NB * NB * NB;
    UpdOp<T> unone{};
    TrsmOp<T> tnone{};
    PotrfOpT<T> pnone{};
    TTGC_HIP(launch_prepare_t<T>(s_main, N, NB, tiles, info));
    const bool la = (flags & TTG_CHOL_F_LOOKAHEAD) != 0;
    const bool rsv = la && s_diag;  // POTRF on its reserved CUs
    hipStream_t sp = la ? s_look : s_main;
    hipStream_t sd = rsv ? s_diag : sp;
    if (la) TTGC_HIP(hipEventRecord(ev_col, s_main));
    auto upd = [&](int off, int cnt, int nsyrk) -> int {
      const int slot = prof ? prof_begin(s_main) : -1;
      if (lowp)
        TTGC_HIP(launch_update_f16x2(s_main, f16_ops() + off, cnt, NB, info, lowp));
      else
        TTGC_HIP(launch_update<T>(s_main, upd_ops() + off, cnt, unone, NB, NB, m1, p1, fast, info));
      if (prof) prof_end(slot, s_main, PROF_UPDATE, (cnt - nsyrk) * 2.0 * b3 + nsyrk * (double)NB * NB * (NB + 1));
      return TTG_CHOL_OK;
    };
    auto potrf = [&](int k) -> int {
      const int slot = prof ? prof_begin(sd) : -1;
      TTGC_HIP(launch_potrf<T>(sd, potrf_ops() + k, 1, pnone, info, potrf_in_ts()));
      if (prof) prof_end(slot, sd, PROF_POTRF, b3 / 3.0);
      if (rsv) {  // TRSM(k) follows on the panel stream
        TTGC_HIP(hipEventRecord(ev_diag, s_diag));
        TTGC_HIP(hipStreamWaitEvent(s_look, ev_diag, 0));
      }
      return TTG_CHOL_OK;
    };
    // Step k: TRSM(k) (+ split) on the panel stream, updA(k) (column k+1),
    // then POTRF(k+1), which needs only updA(k).  On shared CUs updB(k) is
    // released only after POTRF(k+1): the single-workgroup POTRF then finds
    // free CUs instead of waiting behind thousands of update workgroups.  On
    // reserved CUs the two run side by side.  TRSM(k+1) overlaps updB(k).
    if (la) TTGC_HIP(hipStreamWaitEvent(sd, ev_col, 0));
    if (int rc = potrf(0)) return rc;
    for (int k = 0; k < NT; ++k) {
      if (trsm_cnt[k]) {
        const int slot = prof ? prof_begin(sp) : -1;
        TTGC_HIP(launch_trsm<T>(sp, trsm_ops() + trsm_off[k], trsm_cnt[k], tnone, NB, info));
        if (prof) prof_end(slot, sp, PROF_TRSM, trsm_cnt[k] * b3);
        if (lowp) TTGC_HIP(launch_split(sp, split_ops() + split_off[k], split_cnt[k], NB, lowp));
      }
      if (la) {
        TTGC_HIP(hipEventRecord(ev_panel, s_look));
        TTGC_HIP(hipStreamWaitEvent(s_main, ev_panel, 0));
      }
      if (updA_cnt[k]) {
        int rc = upd(updA_off[k], updA_cnt[k], 1);
        if (rc) return rc;
      }
      if (k + 1 < NT) {
        if (la) {
          TTGC_HIP(hipEventRecord(ev_col, s_main));
          TTGC_HIP(hipStreamWaitEvent(sd, ev_col, 0));
        }
        if (int rc = potrf(k + 1)) return rc;
        if (la && !rsv && updB_cnt[k]) {
          TTGC_HIP(hipEventRecord(ev_diag, s_look));
          TTGC_HIP(hipStreamWaitEvent(s_main, ev_diag, 0));
        }
      }
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
    int rc = bind_info(info);
    if (rc) return rc;
    const bool prof = (flags & TTG_CHOL_F_PROFILE) != 0 || prof_enabled();
    if ((flags & TTG_CHOL_F_GRAPH) && !prof && !s_diag) {
      if (pre_bulk) {  // never captured: it runs before the replay, on the caller's stream
        std::function<int(hipStream_t)> hook;
        hook.swap(pre_bulk);
        rc = hook(user);
        hook.swap(pre_bulk);
        if (rc) return rc;
      }
      if (!gexec || bound_info != info) {
        std::function<int(hipStream_t)> hook;  // empty while the schedule is captured
        hook.swap(pre_bulk);
        if (gexec) {
          TTGC_HIP(hipGraphExecDestroy(gexec));
          gexec = nullptr;
        }
        if (graph) {
          TTGC_HIP(hipGraphDestroy(graph));
          graph = nullptr;
        }
        TTGC_HIP(hipStreamBeginCapture(s_main, hipStreamCaptureModeThreadLocal));
        rc = enqueue(info, false);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s_main, &g);
        hook.swap(pre_bulk);
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
    rc = enqueue(info, prof);
    if (rc != TTG_CHOL_OK) return rc;
    TTGC_HIP(hipEventRecord(ev_out, s_main));
    TTGC_HIP(hipStreamWaitEvent(user, ev_out, 0));
    return TTG_CHOL_OK;
  }
};

}  // namespace ttgc
