// extern "C" entry points of libttg_chol: per-tile operations, launch
// grouping and the tiled FP64 driver.
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
#include "chol_overlap.hpp"
#include "chol_tileops.hpp"

namespace ttgc {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}
void clear_error() { g_err[0] = 0; }

constexpr int kDinvBlock = 32 * 32;  // doubles per 32 x 32 inverse block

static int dinv_doubles(int n) { return ((n + 31) / 32) * kDinvBlock; }

// ---------------------------------------------------------------------------
// per-stream library workspaces (descriptor staging and TRSM inverses)
// ---------------------------------------------------------------------------
struct StreamWork {
  DevBuf ops;   // staged descriptors of grouped launches
  DevBuf dinv;  // inverses of diagonal blocks for stand-alone TRSMs
  // pinned host staging of the descriptors; `copied` marks when the last
  // asynchronous copy out of it has finished
  char *host = nullptr;
  size_t host_bytes = 0;
  hipEvent_t copied = nullptr;
  bool pending = false;
  hipError_t host_reserve(size_t n) {
    if (pending) {
      hipError_t e = hipEventSynchronize(copied);
      if (e != hipSuccess) return e;
      pending = false;
    }
    if (!copied) {
      hipError_t e = hipEventCreateWithFlags(&copied, hipEventDisableTiming);
      if (e != hipSuccess) return e;
    }
    if (n <= host_bytes) return hipSuccess;
    if (host) {
      hipError_t e = hipHostFree(host);
      if (e != hipSuccess) return e;
      host = nullptr;
      host_bytes = 0;
    }
    size_t want = n < 65536 ? 65536 : 2 * n;
    hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&host), want, hipHostMallocDefault);
    if (e == hipSuccess) host_bytes = want;
    return e;
  }
};
static std::mutex g_ws_mu;
static std::map<hipStream_t, std::unique_ptr<StreamWork>> g_ws;

static StreamWork &stream_work(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  auto &p = g_ws[s];
  if (!p) p.reset(new StreamWork);
  return *p;
}

// ---------------------------------------------------------------------------
// scratch of the allocation-free TRSM (ttg_chol_dtrsm_rlt_noalloc): the
// inverses of the 32 x 32 diagonal blocks of one L of order <= 512, 128 KiB.
// Every GPU the library runs on holds its own copy until the process ends.
// Host code never uses the array's name as an address: the device address of
// each GPU's copy comes from hipGetSymbolAddress, taken once per device with
// that device current.
// ---------------------------------------------------------------------------
constexpr int kScratchDoubles = (TTG_CHOL_MAX_TILE / 32) * kDinvBlock;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices] = {};

static int scratch_ptr(int dev, double **out) {
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("dtrsm_rlt_noalloc: device %d beyond the %d the library tracks", dev, kMaxDevices);
    return TTG_CHOL_ENOTSUP;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "dtrsm_rlt_noalloc: hipGetSymbolAddress(scratch)");
    if (!p) {
      set_error("dtrsm_rlt_noalloc: hipGetSymbolAddress returned no address");
      return TTG_CHOL_EHIP;
    }
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return TTG_CHOL_OK;
}

// ---------------------------------------------------------------------------
// launch grouping state (per calling thread)
// ---------------------------------------------------------------------------
struct UpdKey {
  double alpha, beta;
  bool ta, tb;
  bool operator<(const UpdKey &o) const { return std::tie(alpha, beta, ta, tb) < std::tie(o.alpha, o.beta, o.ta, o.tb); }
};

// Pointers of a recorded op that point into the group's work area hold the
// byte offset there until the group runs; `rel` bits name them.
constexpr uint8_t kRelA = 1, kRelB = 2;

struct Group {
  bool active = false;
  hipStream_t stream = nullptr;
  std::vector<PotrfOp> potrf;
  // TRSMs B := B L^{-T} (trans) and B := B L^{-1} (notrans), each with the
  // job that forms the inverses of L's 32 x 32 diagonal blocks; rel kRelB:
  // B is a work tile
  std::vector<TrsmOp<double>> trsm, trsm_n;
  std::vector<PotrfOp> trsm_dinv, trsm_n_dinv;
  std::vector<uint8_t> trsm_rel, trsm_n_rel;
  std::map<UpdKey, std::vector<UpdOp<double>>> upd;
  std::map<UpdKey, std::vector<uint8_t>> upd_rel;  // kRelA / kRelB: A / B in the work area
  // passes of the POINV tile operations: before the TRSMs and updates, and
  // after them (rel kRelA: src, kRelB: dst in the work area)
  std::vector<TopOp> pre, post;
  std::vector<uint8_t> pre_rel, post_rel;
  std::vector<DiagZeroOp> zero;
  // FP32 tiles (the graph over float tiles): POTRF (its FP64 workspace pw in
  // the work area), TRSM B := B L^{-T} with its inverse-block jobs, updates
  std::vector<PotrfOpT<float>> potrf_f, trsm_f_dinv;
  std::vector<TrsmOp<float>> trsm_f;
  std::map<UpdKey, std::vector<UpdOp<float>>> upd_f;
  size_t work = 0;  // bytes of work tiles
  struct Norm {
    int n;
    const double *x;
    double *r;
  };
  std::vector<Norm> norm_after;  // norms of memory the group writes
  struct Write {
    const double *p;
    int m, n, ld;
  };
  std::vector<Write> writes;  // outputs of the POINV tile operations
  std::vector<PlgsyOp> plgsy;
  Footprints fp;       // what the recorded calls read and write
  int last_launches = 0;  // launches of the last group (all its parts)
  int splits = 0;         // parts run early because a call depended on an earlier one

  void clear() {
    potrf.clear();
    trsm.clear();
    trsm_n.clear();
    trsm_dinv.clear();
    trsm_n_dinv.clear();
    trsm_rel.clear();
    trsm_n_rel.clear();
    upd.clear();
    upd_rel.clear();
    pre.clear();
    post.clear();
    pre_rel.clear();
    post_rel.clear();
    zero.clear();
    potrf_f.clear();
    trsm_f_dinv.clear();
    trsm_f.clear();
    upd_f.clear();
    work = 0;
    norm_after.clear();
    writes.clear();
    plgsy.clear();
    fp.clear();
  }
  // a work tile of m x n doubles (leading dimension m), as an offset
  double *work_tile(int m, int n) {
    const size_t off = work;
    work += ((size_t)m * n * sizeof(double) + 255) & ~size_t(255);
    return reinterpret_cast<double *>(off);
  }
  void add_upd(const UpdKey &k, const UpdOp<double> &op, uint8_t rel) {
    upd[k].push_back(op);
    upd_rel[k].push_back(rel);
  }
};
static thread_local Group g_group;

static bool grouping(hipStream_t s) { return g_group.active && g_group.stream == s; }

// [p, p + elements of an m x n column-major block with leading dimension ld)
static bool overlaps(const double *x, int nx, const double *p, int m, int n, int ld) {
  if (nx <= 0 || m <= 0 || n <= 0) return false;
  const double *end = p + (long)(n - 1) * ld + m;
  return x < end && p < x + nx;
}

bool group_defer_nrm2(hipStream_t s, int n, const double *x, double *r) {
  if (!grouping(s)) return false;
  Group &g = g_group;
  bool written = false;
  for (auto &o : g.plgsy) written = written || overlaps(x, n, o.A, o.m, o.n, (int)o.lda);
  for (auto &o : g.potrf) written = written || overlaps(x, n, o.A, o.n, o.n, o.lda);
  for (size_t i = 0; i < g.trsm.size(); ++i)
    written = written || (!(g.trsm_rel[i] & kRelB) && overlaps(x, n, g.trsm[i].B, g.trsm[i].m, g.trsm[i].n, g.trsm[i].ldb));
  for (auto &kv : g.upd)
    for (auto &o : kv.second) written = written || overlaps(x, n, o.C, o.m, o.n, o.ldc);
  for (auto &w : g.writes) written = written || overlaps(x, n, w.p, w.m, w.n, w.ld);
  if (written) g.norm_after.push_back(Group::Norm{n, x, r});
  return written;
}

// ---------------------------------------------------------------------------
// POINV tile operations as passes of a group (work tiles are offsets into the
// group's work area until it runs)
// ---------------------------------------------------------------------------
static void rec_trsm(Group &g, bool trans, const double *L, int ldl, int order, double *B, int m, int ldb,
                     uint8_t rel) {
  TrsmOp<double> op{L, B, nullptr, m, order, ldl, ldb};
  PotrfOp dj{const_cast<double *>(L), nullptr, nullptr, order, ldl, 0, 0};
  if (trans) {
    g.trsm.push_back(op);
    g.trsm_dinv.push_back(dj);
    g.trsm_rel.push_back(rel);
  } else {
    g.trsm_n.push_back(op);
    g.trsm_n_dinv.push_back(dj);
    g.trsm_n_rel.push_back(rel);
  }
}

static void rec_pass(std::vector<TopOp> &v, std::vector<uint8_t> &rel, const TopOp &op, uint8_t r) {
  v.push_back(op);
  rel.push_back(r);
}

// A := A^{-1} (lower, non-unit): W := I; W := W L^{-T} = L^{-T}; lower(A) := lower(W^T)
static void rec_trtri(Group &g, int n, double *A, int lda, int *d_info, int info_offset) {
  if (d_info) g.zero.push_back(DiagZeroOp{A, d_info, n, lda, info_offset});
  double *W = g.work_tile(n, n);
  rec_pass(g.pre, g.pre_rel, TopOp{nullptr, W, n, n, n, n, kTopIdentity, 1.0}, kRelB);
  rec_trsm(g, true, A, lda, n, W, n, n, kRelB);
  rec_pass(g.post, g.post_rel, TopOp{W, A, n, n, n, lda, kTopTransLower, 1.0}, kRelA);
  g.writes.push_back(Group::Write{A, n, n, lda});
}

// B (m x n) := alpha L^{-1} B = alpha (B^T L^{-T})^T, L m x m
static void rec_trsm_lln(Group &g, int m, int n, double alpha, const double *L, int ldl, double *B, int ldb) {
  double *W = g.work_tile(n, m);
  rec_pass(g.pre, g.pre_rel, TopOp{B, W, m, n, ldb, n, kTopTrans, 1.0}, kRelB);
  rec_trsm(g, true, L, ldl, m, W, n, n, kRelB);
  rec_pass(g.post, g.post_rel, TopOp{W, B, n, m, n, ldb, kTopTrans, alpha}, kRelA);
  g.writes.push_back(Group::Write{B, m, n, ldb});
}

// B (m x n) := alpha B L^{-1}, L n x n
static void rec_trsm_rln(Group &g, int m, int n, double alpha, const double *L, int ldl, double *B, int ldb) {
  rec_trsm(g, false, L, ldl, n, B, m, ldb, 0);
  if (alpha != 1.0) rec_pass(g.post, g.post_rel, TopOp{B, B, m, n, ldb, ldb, kTopCopy, alpha}, 0);
  g.writes.push_back(Group::Write{B, m, n, ldb});
}

// lower(A) := L^T L (L the lower triangle of A): W := lower(A); lower(A) := W^T W
static void rec_lauum(Group &g, int n, double *A, int lda) {
  double *W = g.work_tile(n, n);
  rec_pass(g.pre, g.pre_rel, TopOp{A, W, n, n, lda, n, kTopLowerZero, 1.0}, kRelB);
  g.add_upd(UpdKey{1.0, 0.0, true, true}, UpdOp<double>{W, W, A, n, n, n, n, n, lda, 1}, kRelA | kRelB);
  g.writes.push_back(Group::Write{A, n, n, lda});
}

// B (m x n) := alpha L^T B, L m x m lower: W1 := lower(L), W2 := B; B := alpha W1^T W2
static void rec_trmm_llt(Group &g, int m, int n, double alpha, const double *L, int ldl, double *B, int ldb) {
  double *W1 = g.work_tile(m, m), *W2 = g.work_tile(m, n);
  rec_pass(g.pre, g.pre_rel, TopOp{L, W1, m, m, ldl, m, kTopLowerZero, 1.0}, kRelB);
  rec_pass(g.pre, g.pre_rel, TopOp{B, W2, m, n, ldb, m, kTopCopy, 1.0}, kRelB);
  g.add_upd(UpdKey{alpha, 0.0, true, true}, UpdOp<double>{W1, W2, B, m, n, m, m, m, ldb, 0}, kRelA | kRelB);
  g.writes.push_back(Group::Write{B, m, n, ldb});
}

static int run_group(Group &g, hipStream_t s);

// a group of its own for a call outside any open group
template <typename F>
static int run_alone(hipStream_t s, F &&record) {
  Group local;
  record(local);
  return run_group(local, s);
}

// Run the recorded work of a group on stream s: every descriptor array and
// the work tiles live in one staging buffer of the stream, filled by one
// copy.  The ops of a group are independent of each other (the tasks of one
// scheduling round), so the kinds run in a fixed order; the passes of one
// POINV operation are ordered by it: diagonal checks, pre passes, TRSM
// inverses, TRSMs, updates, post passes, then deferred norms.
static int run_group(Group &g, hipStream_t s) {
  StreamWork &w = stream_work(s);
  size_t off = 0;
  auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
  auto place = [&](size_t bytes) {
    const size_t o = off;
    off = align(off + bytes);
    return o;
  };
  const size_t o_plgsy = place(g.plgsy.size() * sizeof(PlgsyOp));
  const size_t o_potrf = place(g.potrf.size() * sizeof(PotrfOp));
  const size_t o_zero = place(g.zero.size() * sizeof(DiagZeroOp));
  const size_t o_pre = place(g.pre.size() * sizeof(TopOp));
  const size_t o_post = place(g.post.size() * sizeof(TopOp));
  const size_t ndinv = g.trsm_dinv.size() + g.trsm_n_dinv.size();
  const size_t o_dinvops = place(ndinv * sizeof(PotrfOp));
  const size_t o_trsm = place(g.trsm.size() * sizeof(TrsmOp<double>));
  const size_t o_trsm_n = place(g.trsm_n.size() * sizeof(TrsmOp<double>));
  std::vector<size_t> o_upd;
  for (auto &kv : g.upd) o_upd.push_back(place(kv.second.size() * sizeof(UpdOp<double>)));
  const size_t o_potrf_f = place(g.potrf_f.size() * sizeof(PotrfOpT<float>));
  const size_t o_dinvops_f = place(g.trsm_f_dinv.size() * sizeof(PotrfOpT<float>));
  const size_t o_trsm_f = place(g.trsm_f.size() * sizeof(TrsmOp<float>));
  std::vector<size_t> o_upd_f;
  for (auto &kv : g.upd_f) o_upd_f.push_back(place(kv.second.size() * sizeof(UpdOp<float>)));
  const size_t desc_bytes = off;  // descriptors end here: copied from the host
  size_t dinv_bytes = 0;
  for (auto &d : g.trsm_dinv) dinv_bytes += sizeof(double) * dinv_doubles(d.n);
  for (auto &d : g.trsm_n_dinv) dinv_bytes += sizeof(double) * dinv_doubles(d.n);
  for (auto &d : g.trsm_f_dinv) dinv_bytes += sizeof(float) * dinv_doubles(d.n);
  const size_t dinv_off = place(dinv_bytes);
  const size_t work_off = place(g.work);
  if (off == 0) return TTG_CHOL_OK;
  TTGC_HIP(w.ops.reserve(off));
  char *base = static_cast<char *>(w.ops.ptr);
  char *work = base + work_off;
  auto reloc = [&](const double *p) {
    return reinterpret_cast<double *>(work + reinterpret_cast<uintptr_t>(p));
  };
  // inverse-block workspace of the TRSMs and the work tiles
  {
    size_t p = dinv_off;
    for (size_t i = 0; i < g.trsm.size(); ++i) {
      g.trsm_dinv[i].dinv = reinterpret_cast<double *>(base + p);
      g.trsm[i].dinv = g.trsm_dinv[i].dinv;
      if (g.trsm_rel[i] & kRelB) g.trsm[i].B = reloc(g.trsm[i].B);
      p += sizeof(double) * dinv_doubles(g.trsm_dinv[i].n);
    }
    for (size_t i = 0; i < g.trsm_n.size(); ++i) {
      g.trsm_n_dinv[i].dinv = reinterpret_cast<double *>(base + p);
      g.trsm_n[i].dinv = g.trsm_n_dinv[i].dinv;
      if (g.trsm_n_rel[i] & kRelB) g.trsm_n[i].B = reloc(g.trsm_n[i].B);
      p += sizeof(double) * dinv_doubles(g.trsm_n_dinv[i].n);
    }
    for (size_t i = 0; i < g.trsm_f.size(); ++i) {
      g.trsm_f_dinv[i].dinv = reinterpret_cast<float *>(base + p);
      g.trsm_f[i].dinv = g.trsm_f_dinv[i].dinv;
      p += sizeof(float) * dinv_doubles(g.trsm_f_dinv[i].n);
    }
  }
  for (auto &o : g.potrf_f) o.pw = reloc(o.pw);
  for (size_t i = 0; i < g.pre.size(); ++i) {
    if (g.pre_rel[i] & kRelA) g.pre[i].src = reloc(g.pre[i].src);
    if (g.pre_rel[i] & kRelB) g.pre[i].dst = reloc(g.pre[i].dst);
  }
  for (size_t i = 0; i < g.post.size(); ++i) {
    if (g.post_rel[i] & kRelA) g.post[i].src = reloc(g.post[i].src);
    if (g.post_rel[i] & kRelB) g.post[i].dst = reloc(g.post[i].dst);
  }
  for (auto &kv : g.upd) {
    auto &rel = g.upd_rel[kv.first];
    for (size_t i = 0; i < kv.second.size(); ++i) {
      if (rel[i] & kRelA) kv.second[i].A = reloc(kv.second[i].A);
      if (rel[i] & kRelB) kv.second[i].B = reloc(kv.second[i].B);
    }
  }
  TTGC_HIP(w.host_reserve(desc_bytes));
  char *host = w.host;
  auto put = [&](size_t o, const void *src, size_t bytes) {
    if (bytes) memcpy(host + o, src, bytes);
  };
  put(o_plgsy, g.plgsy.data(), g.plgsy.size() * sizeof(PlgsyOp));
  put(o_potrf, g.potrf.data(), g.potrf.size() * sizeof(PotrfOp));
  put(o_zero, g.zero.data(), g.zero.size() * sizeof(DiagZeroOp));
  put(o_pre, g.pre.data(), g.pre.size() * sizeof(TopOp));
  put(o_post, g.post.data(), g.post.size() * sizeof(TopOp));
  put(o_dinvops, g.trsm_dinv.data(), g.trsm_dinv.size() * sizeof(PotrfOp));
  put(o_dinvops + g.trsm_dinv.size() * sizeof(PotrfOp), g.trsm_n_dinv.data(), g.trsm_n_dinv.size() * sizeof(PotrfOp));
  put(o_trsm, g.trsm.data(), g.trsm.size() * sizeof(TrsmOp<double>));
  put(o_trsm_n, g.trsm_n.data(), g.trsm_n.size() * sizeof(TrsmOp<double>));
  {
    size_t i = 0;
    for (auto &kv : g.upd) put(o_upd[i++], kv.second.data(), kv.second.size() * sizeof(UpdOp<double>));
  }
  put(o_potrf_f, g.potrf_f.data(), g.potrf_f.size() * sizeof(PotrfOpT<float>));
  put(o_dinvops_f, g.trsm_f_dinv.data(), g.trsm_f_dinv.size() * sizeof(PotrfOpT<float>));
  put(o_trsm_f, g.trsm_f.data(), g.trsm_f.size() * sizeof(TrsmOp<float>));
  {
    size_t i = 0;
    for (auto &kv : g.upd_f) put(o_upd_f[i++], kv.second.data(), kv.second.size() * sizeof(UpdOp<float>));
  }
  if (desc_bytes) {
    TTGC_HIP(hipMemcpyAsync(base, host, desc_bytes, hipMemcpyHostToDevice, s));
    TTGC_HIP(hipEventRecord(w.copied, s));
    w.pending = true;
  }
  auto max_dims = [](const std::vector<TopOp> &v, int &mm, int &mn) {
    mm = mn = 0;
    for (auto &o : v) {
      mm = std::max(mm, o.m);
      mn = std::max(mn, o.n);
    }
  };
  int launches = 0;
  if (!g.plgsy.empty()) {
    int max_n = 0;
    for (auto &o : g.plgsy) max_n = std::max(max_n, o.n);
    TTGC_HIP(launch_plgsy_batched(s, reinterpret_cast<PlgsyOp *>(base + o_plgsy), (int)g.plgsy.size(), max_n));
    launches += ((int)g.plgsy.size() + 65534) / 65535;
  }
  PotrfOp none{};
  if (!g.potrf.empty()) {
    TTGC_HIP(launch_potrf(s, reinterpret_cast<PotrfOp *>(base + o_potrf), (int)g.potrf.size(), none, nullptr));
    ++launches;
  }
  if (!g.zero.empty()) {
    TTGC_HIP(launch_diag_zero(s, reinterpret_cast<DiagZeroOp *>(base + o_zero), (int)g.zero.size()));
    ++launches;
  }
  if (!g.pre.empty()) {
    int mm, mn;
    max_dims(g.pre, mm, mn);
    TTGC_HIP(launch_tile_ops(s, reinterpret_cast<TopOp *>(base + o_pre), (int)g.pre.size(), mm, mn));
    ++launches;
  }
  if (ndinv) {
    TTGC_HIP(launch_dinv(s, reinterpret_cast<PotrfOp *>(base + o_dinvops), (int)ndinv, none));
    ++launches;
  }
  TrsmOp<double> tnone{};
  if (!g.trsm.empty()) {
    int max_m = 0;
    for (auto &t : g.trsm) max_m = std::max(max_m, t.m);
    TTGC_HIP(launch_trsm<double>(s, reinterpret_cast<TrsmOp<double> *>(base + o_trsm), (int)g.trsm.size(), tnone,
                                 max_m, nullptr, true));
    ++launches;
  }
  if (!g.trsm_n.empty()) {
    int max_m = 0;
    for (auto &t : g.trsm_n) max_m = std::max(max_m, t.m);
    TTGC_HIP(launch_trsm<double>(s, reinterpret_cast<TrsmOp<double> *>(base + o_trsm_n), (int)g.trsm_n.size(),
                                 tnone, max_m, nullptr, false));
    ++launches;
  }
  {
    size_t i = 0;
    for (auto &kv : g.upd) {
      int max_m = 0, max_n = 0;
      bool fast = true;
      for (auto &u : kv.second) {
        max_m = std::max(max_m, u.m);
        max_n = std::max(max_n, u.n);
      }
      for (auto &u : kv.second) fast = fast && update_fast(u, max_m);
      UpdOp<double> unone{};
      TTGC_HIP(launch_update<double>(s, reinterpret_cast<UpdOp<double> *>(base + o_upd[i]), (int)kv.second.size(),
                                     unone, max_m, max_n, kv.first.alpha, kv.first.beta, fast, nullptr, kv.first.tb,
                                     kv.first.ta));
      ++launches;
      ++i;
    }
  }
  // FP32 tiles
  if (!g.potrf_f.empty()) {
    PotrfOpT<float> fnone{};
    TTGC_HIP(launch_potrf<float>(s, reinterpret_cast<PotrfOpT<float> *>(base + o_potrf_f), (int)g.potrf_f.size(),
                                 fnone, nullptr));
    ++launches;
  }
  if (!g.trsm_f.empty()) {
    PotrfOpT<float> fnone{};
    TTGC_HIP(launch_dinv<float>(s, reinterpret_cast<PotrfOpT<float> *>(base + o_dinvops_f), (int)g.trsm_f_dinv.size(),
                                fnone));
    int max_m = 0;
    for (auto &t : g.trsm_f) max_m = std::max(max_m, t.m);
    TrsmOp<float> tfnone{};
    TTGC_HIP(launch_trsm<float>(s, reinterpret_cast<TrsmOp<float> *>(base + o_trsm_f), (int)g.trsm_f.size(), tfnone,
                                max_m, nullptr, true));
    launches += 2;
  }
  {
    size_t i = 0;
    for (auto &kv : g.upd_f) {
      int max_m = 0, max_n = 0;
      bool fast = true;
      for (auto &u : kv.second) {
        max_m = std::max(max_m, u.m);
        max_n = std::max(max_n, u.n);
      }
      for (auto &u : kv.second) fast = fast && update_fast(u, max_m);
      UpdOp<float> unone{};
      TTGC_HIP(launch_update<float>(s, reinterpret_cast<UpdOp<float> *>(base + o_upd_f[i]), (int)kv.second.size(),
                                    unone, max_m, max_n, (float)kv.first.alpha, (float)kv.first.beta, fast, nullptr,
                                    false, false));
      ++launches;
      ++i;
    }
  }
  if (!g.post.empty()) {
    int mm, mn;
    max_dims(g.post, mm, mn);
    TTGC_HIP(launch_tile_ops(s, reinterpret_cast<TopOp *>(base + o_post), (int)g.post.size(), mm, mn));
    ++launches;
  }
  for (auto &nm : g.norm_after) TTGC_HIP(launch_nrm2(s, nm.n, nm.x, nm.r));
  g.last_launches += launches;
  return TTG_CHOL_OK;
}

// A call about to be recorded into the open group on s, with the given reads
// and writes: if it reads what a recorded call writes, or writes what one
// reads or writes (a task body that issues dependent calls, e.g. SYRK and
// then POTRF on one tile), the group so far runs first, so that stream order
// keeps the two in program order, and recording goes on in an empty group.
template <size_t NI, size_t NO>
static int group_admit(hipStream_t s, const Span (&ins)[NI], const Span (&outs)[NO]) {
  Group &g = g_group;
  if (g.fp.conflicts(ins, outs)) {
    ++g.splits;
    const int rc = run_group(g, s);
    g.clear();
    if (rc != TTG_CHOL_OK) return rc;
  }
  g.fp.add(ins, outs);
  return TTG_CHOL_OK;
}
template <typename T>
static Span blk(const T *p, long m, long n, long ld) {
  return block_span(p, m, n, ld, sizeof(T));
}

bool group_record_plgsy(hipStream_t s, const PlgsyOp &op) {
  if (!grouping(s)) return false;
  const Span none[1], outs[1] = {block_span(op.A, op.m, op.n, op.lda, sizeof(double))};
  if (group_admit(s, none, outs) != TTG_CHOL_OK) return false;  // the caller launches it alone
  g_group.plgsy.push_back(op);
  return true;
}

}  // namespace ttgc

using namespace ttgc;

namespace {
// process-wide schedule knobs; -1: not set (environment or default)
int g_knobs[TTG_CHOL_KNOB_COUNT] = {-1, -1, -1, -1, -1};
std::mutex g_knob_mu;
}  // namespace

int ttgc::plan_knob(int which, const char *env, int def) {
  {
    std::lock_guard<std::mutex> g(g_knob_mu);
    if (which >= 0 && which < TTG_CHOL_KNOB_COUNT && g_knobs[which] >= 0) return g_knobs[which];
  }
  if (const char *e = env ? getenv(env) : nullptr) return atoi(e);
  return def;
}

extern "C" {

const char *ttg_chol_version(void) { return "ttg_chol 0.1 (gfx950)"; }

const char *ttg_chol_last_error(void) { return ttgc::g_err; }

int ttg_chol_num_devices(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int ttg_chol_set_knob(int knob, int value) {
  clear_error();
  static const int lo[TTG_CHOL_KNOB_COUNT] = {2, 0, 1, 0, 0}, hi[TTG_CHOL_KNOB_COUNT] = {8, 64, 8, 1, 8};
  if (knob < 0 || knob >= TTG_CHOL_KNOB_COUNT || (value >= 0 && (value < lo[knob] || value > hi[knob]))) {
    set_error("set_knob: knob %d cannot take %d", knob, value);
    return TTG_CHOL_EINVAL;
  }
  std::lock_guard<std::mutex> g(g_knob_mu);
  g_knobs[knob] = value < 0 ? -1 : value;
  return TTG_CHOL_OK;
}

int ttg_chol_get_knob(int knob) {
  clear_error();
  switch (knob) {
    case TTG_CHOL_KNOB_GROUP: return plan_knob(knob, "TTG_CHOL_GROUP", 3);
    case TTG_CHOL_KNOB_POTRF_CUS: return plan_knob(knob, "TTG_CHOL_POTRF_CUS", 8);
    case TTG_CHOL_KNOB_TILE_GROUP: return plan_knob(knob, "TTG_CHOL_TILE_GROUP", 2);
    case TTG_CHOL_KNOB_PAIRED: return plan_knob(knob, "TTG_CHOL_PAIRED", 1);
    case TTG_CHOL_KNOB_SWEEP: return plan_knob(knob, "TTG_CHOL_SWEEP", 2);
    default: set_error("get_knob: no knob %d", knob); return TTG_CHOL_EINVAL;
  }
}

int ttg_chol_release_plans(void) {
  clear_error();
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "release_plans: hipDeviceSynchronize");
  release_mixed_plans();  // before the solve plans: a mixed plan's solve plan lives there
  release_poinv_plans();
  release_solve_plans();
  release_tiled_plans();
  release_check_work();
  return TTG_CHOL_OK;
}

int ttg_chol_dpotrf_tile(void *stream, int n, double *A, int lda, int *d_info, int info_offset) {
  clear_error();
  if (n < 0 || n > TTG_CHOL_MAX_TILE || lda < std::max(1, n) || (n > 0 && !A)) {
    set_error("dpotrf_tile: bad argument (n=%d lda=%d)", n, lda);
    return TTG_CHOL_EINVAL;
  }
  hipStream_t s = as_stream(stream);
  PotrfOp op{A, nullptr, d_info, n, lda, info_offset, 1};
  if (n == 0) {
    if (!d_info) return TTG_CHOL_OK;
    op.n = 0;  // the kernel stores 0 into *d_info
  }
  if (grouping(s)) {
    const Span none[1], outs[1] = {blk(A, n, n, lda)};
    if (int rc = group_admit(s, none, outs)) return rc;
    g_group.potrf.push_back(op);
    return TTG_CHOL_OK;
  }
  TTGC_HIP(launch_potrf<double>(s, nullptr, 1, op, nullptr));
  return TTG_CHOL_OK;
}

int ttg_chol_dtrsm_rlt(void *stream, int m, int n, const double *L, int ldl, double *B, int ldb) {
  clear_error();
  if (m < 0 || n < 0 || n > TTG_CHOL_MAX_TILE || ldl < std::max(1, n) || ldb < std::max(1, m) ||
      (m > 0 && n > 0 && (!L || !B))) {
    set_error("dtrsm_rlt: bad argument (m=%d n=%d ldl=%d ldb=%d)", m, n, ldl, ldb);
    return TTG_CHOL_EINVAL;
  }
  if (m == 0 || n == 0) return TTG_CHOL_OK;
  hipStream_t s = as_stream(stream);
  TrsmOp<double> op{L, B, nullptr, m, n, ldl, ldb};
  PotrfOp dj{const_cast<double *>(L), nullptr, nullptr, n, ldl, 0, 0};
  if (grouping(s)) {
    const Span ins[1] = {blk(L, n, n, ldl)}, outs[1] = {blk(B, m, n, ldb)};
    if (int rc = group_admit(s, ins, outs)) return rc;
    g_group.trsm.push_back(op);
    g_group.trsm_dinv.push_back(dj);
    g_group.trsm_rel.push_back(0);
    return TTG_CHOL_OK;
  }
  StreamWork &w = stream_work(s);
  TTGC_HIP(w.dinv.reserve(sizeof(double) * dinv_doubles(n)));
  dj.dinv = static_cast<double *>(w.dinv.ptr);
  op.dinv = dj.dinv;
  TTGC_HIP(launch_dinv<double>(s, nullptr, 1, dj));
  TTGC_HIP(launch_trsm<double>(s, nullptr, 1, op, m, nullptr));
  return TTG_CHOL_OK;
}

int ttg_chol_dtrsm_rlt_noalloc(void *stream, int m, int n, const double *L, int ldl, double *B, int ldb) {
  clear_error();
  if (n > TTG_CHOL_MAX_TILE) {
    set_error("dtrsm_rlt_noalloc: n=%d exceeds %d, the order its fixed scratch (%d doubles) serves", n,
              TTG_CHOL_MAX_TILE, kScratchDoubles);
    return TTG_CHOL_EINVAL;
  }
  if (m < 0 || n < 0 || ldl < std::max(1, n) || ldb < std::max(1, m) || (m > 0 && n > 0 && (!L || !B))) {
    set_error("dtrsm_rlt_noalloc: bad argument (m=%d n=%d ldl=%d ldb=%d)", m, n, ldl, ldb);
    return TTG_CHOL_EINVAL;
  }
  if (m == 0 || n == 0) return TTG_CHOL_OK;
  hipStream_t s = as_stream(stream);
  if (grouping(s)) {
    // a group stages its descriptors in library-allocated memory
    set_error("dtrsm_rlt_noalloc: a launch group is open on this stream; use ttg_chol_dtrsm_rlt there");
    return TTG_CHOL_ESTATE;
  }
  int dev = 0;
  TTGC_HIP(hipGetDevice(&dev));
  if (s) {
    int sdev = 0;
    TTGC_HIP(hipStreamGetDevice(s, &sdev));
    if (sdev != dev) {
      set_error("dtrsm_rlt_noalloc: the stream belongs to device %d, the current device is %d", sdev, dev);
      return TTG_CHOL_EINVAL;
    }
  }
  double *dinv = nullptr;
  int rc = scratch_ptr(dev, &dinv);
  if (rc != TTG_CHOL_OK) return rc;
  // dinv_kernel writes every block inverse the TRSM reads (padding blocks as
  // identity), so the scratch needs no clearing between calls
  PotrfOp dj{const_cast<double *>(L), dinv, nullptr, n, ldl, 0, 0};
  TrsmOp<double> op{L, B, dinv, m, n, ldl, ldb};
  TTGC_HIP(launch_dinv<double>(s, nullptr, 1, dj));
  TTGC_HIP(launch_trsm<double>(s, nullptr, 1, op, m, nullptr));
  return TTG_CHOL_OK;
}

static int update_entry(const char *name, hipStream_t s, int m, int n, int k, double alpha, const double *A,
                        int lda, const double *B, int ldb, double beta, double *C, int ldc, int lower) {
  if (m < 0 || n < 0 || k < 0 || lda < std::max(1, m) || ldb < std::max(1, n) || ldc < std::max(1, m) ||
      (m > 0 && n > 0 && (!C || (k > 0 && (!A || !B))))) {
    set_error("%s: bad argument (m=%d n=%d k=%d lda=%d ldb=%d ldc=%d)", name, m, n, k, lda, ldb, ldc);
    return TTG_CHOL_EINVAL;
  }
  if (m == 0 || n == 0) return TTG_CHOL_OK;
  if (k == 0) {
    // C := beta C ; expressed as an update with a zero contribution
    alpha = 0.0;
  }
  UpdOp<double> op{A, B, C, m, n, k, lda, ldb, ldc, lower};
  if (grouping(s)) {
    const Span ins[2] = {blk(A, m, k, lda), blk(B, n, k, ldb)}, outs[1] = {blk(C, m, n, ldc)};
    if (int rc = group_admit(s, ins, outs)) return rc;
    g_group.add_upd(UpdKey{alpha, beta, false, false}, op, 0);
    return TTG_CHOL_OK;
  }
  TTGC_HIP(launch_update<double>(s, nullptr, 1, op, m, n, alpha, beta, update_fast(op, m), nullptr));
  return TTG_CHOL_OK;
}

int ttg_chol_dsyrk_ln(void *stream, int n, int k, double alpha, const double *A, int lda, double beta, double *C,
                      int ldc) {
  clear_error();
  return update_entry("dsyrk_ln", as_stream(stream), n, n, k, alpha, A, lda, A, lda, beta, C, ldc, 1);
}

int ttg_chol_dgemm_nt(void *stream, int m, int n, int k, double alpha, const double *A, int lda, const double *B,
                      int ldb, double beta, double *C, int ldc) {
  clear_error();
  return update_entry("dgemm_nt", as_stream(stream), m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, 0);
}

// ---- FP32 tiles: the potrf graph over float tiles (the reference's bodies
// are templated on the element type, potrf.h:90-631, and call the S
// variants of the vendor routines for float)
int ttg_chol_spotrf_tile(void *stream, int n, float *A, int lda, int *d_info, int info_offset) {
  clear_error();
  if (n < 0 || n > TTG_CHOL_MAX_TILE || lda < std::max(1, n) || (n > 0 && !A)) {
    set_error("spotrf_tile: bad argument (n=%d lda=%d)", n, lda);
    return TTG_CHOL_EINVAL;
  }
  hipStream_t s = as_stream(stream);
  auto rec = [&](Group &g) {
    PotrfOpT<float> op{A, nullptr, d_info, n, lda, info_offset, 1, g.work_tile(std::max(n, 1), std::max(n, 1))};
    g.potrf_f.push_back(op);
  };
  if (n == 0 && !d_info) return TTG_CHOL_OK;
  if (grouping(s)) {
    const Span none[1], outs[1] = {blk(A, n, n, lda)};
    if (int rc = group_admit(s, none, outs)) return rc;
    rec(g_group);
    return TTG_CHOL_OK;
  }
  return run_alone(s, rec);
}

int ttg_chol_strsm_rlt(void *stream, int m, int n, const float *L, int ldl, float *B, int ldb) {
  clear_error();
  if (m < 0 || n < 0 || n > TTG_CHOL_MAX_TILE || ldl < std::max(1, n) || ldb < std::max(1, m) ||
      (m > 0 && n > 0 && (!L || !B))) {
    set_error("strsm_rlt: bad argument (m=%d n=%d ldl=%d ldb=%d)", m, n, ldl, ldb);
    return TTG_CHOL_EINVAL;
  }
  if (m == 0 || n == 0) return TTG_CHOL_OK;
  hipStream_t s = as_stream(stream);
  auto rec = [&](Group &g) {
    g.trsm_f.push_back(TrsmOp<float>{L, B, nullptr, m, n, ldl, ldb});
    g.trsm_f_dinv.push_back(PotrfOpT<float>{const_cast<float *>(L), nullptr, nullptr, n, ldl, 0, 0, nullptr});
  };
  if (grouping(s)) {
    const Span ins[1] = {blk(L, n, n, ldl)}, outs[1] = {blk(B, m, n, ldb)};
    if (int rc = group_admit(s, ins, outs)) return rc;
    rec(g_group);
    return TTG_CHOL_OK;
  }
  return run_alone(s, rec);
}

static int update_f_entry(const char *name, hipStream_t s, int m, int n, int k, float alpha, const float *A, int lda,
                          const float *B, int ldb, float beta, float *C, int ldc, int lower) {
  if (m < 0 || n < 0 || k < 0 || lda < std::max(1, m) || ldb < std::max(1, n) || ldc < std::max(1, m) ||
      (m > 0 && n > 0 && (!C || (k > 0 && (!A || !B))))) {
    set_error("%s: bad argument (m=%d n=%d k=%d lda=%d ldb=%d ldc=%d)", name, m, n, k, lda, ldb, ldc);
    return TTG_CHOL_EINVAL;
  }
  if (m == 0 || n == 0) return TTG_CHOL_OK;
  if (k == 0) alpha = 0.0f;
  const UpdOp<float> op{A, B, C, m, n, k, lda, ldb, ldc, lower};
  const UpdKey key{alpha, beta, false, false};
  if (grouping(s)) {
    const Span ins[2] = {blk(A, m, k, lda), blk(B, n, k, ldb)}, outs[1] = {blk(C, m, n, ldc)};
    if (int rc = group_admit(s, ins, outs)) return rc;
    g_group.upd_f[key].push_back(op);
    return TTG_CHOL_OK;
  }
  return run_alone(s, [&](Group &g) { g.upd_f[key].push_back(op); });
}

int ttg_chol_ssyrk_ln(void *stream, int n, int k, float alpha, const float *A, int lda, float beta, float *C,
                      int ldc) {
  clear_error();
  return update_f_entry("ssyrk_ln", as_stream(stream), n, n, k, alpha, A, lda, A, lda, beta, C, ldc, 1);
}

int ttg_chol_sgemm_nt(void *stream, int m, int n, int k, float alpha, const float *A, int lda, const float *B,
                      int ldb, float beta, float *C, int ldc) {
  clear_error();
  return update_f_entry("sgemm_nt", as_stream(stream), m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, 0);
}

// ---- POINV tile operations (examples/potrf/trtri_L.h, lauum.h)
static bool bad_tile(int n, int ld, const void *p) { return n < 0 || n > TTG_CHOL_MAX_TILE || ld < std::max(1, n) || (n > 0 && !p); }

int ttg_chol_dtrtri_tile(void *stream, int n, double *A, int lda, int *d_info, int info_offset) {
  clear_error();
  if (bad_tile(n, lda, A)) {
    set_error("dtrtri_tile: bad argument (n=%d lda=%d)", n, lda);
    return TTG_CHOL_EINVAL;
  }
  hipStream_t s = as_stream(stream);
  if (n == 0) {  // only *d_info := 0 (by the diagonal check of an empty tile)
    if (!d_info) return TTG_CHOL_OK;
    const DiagZeroOp z{A, d_info, 0, lda, info_offset};
    if (grouping(s)) {
      g_group.zero.push_back(z);
      return TTG_CHOL_OK;
    }
    return run_alone(s, [&](Group &g) { g.zero.push_back(z); });
  }
  if (grouping(s)) {
    const Span none[1], outs[1] = {blk(A, n, n, lda)};
    if (int rc = group_admit(s, none, outs)) return rc;
    rec_trtri(g_group, n, A, lda, d_info, info_offset);
    return TTG_CHOL_OK;
  }
  return run_alone(s, [&](Group &g) { rec_trtri(g, n, A, lda, d_info, info_offset); });
}

static int trsm_args(const char *name, int m, int n, int order, const double *L, int ldl, const double *B, int ldb) {
  if (m < 0 || n < 0 || order > TTG_CHOL_MAX_TILE || m > TTG_CHOL_MAX_TILE || n > TTG_CHOL_MAX_TILE ||
      ldl < std::max(1, order) || ldb < std::max(1, m) || (m > 0 && n > 0 && (!L || !B))) {
    set_error("%s: bad argument (m=%d n=%d ldl=%d ldb=%d)", name, m, n, ldl, ldb);
    return TTG_CHOL_EINVAL;
  }
  return TTG_CHOL_OK;
}

int ttg_chol_dtrsm_lln(void *stream, int m, int n, double alpha, const double *L, int ldl, double *B, int ldb) {
  clear_error();
  if (int rc = trsm_args("dtrsm_lln", m, n, m, L, ldl, B, ldb)) return rc;
  if (m == 0 || n == 0) return TTG_CHOL_OK;
  hipStream_t s = as_stream(stream);
  if (grouping(s)) {
    const Span ins[1] = {blk(L, m, m, ldl)}, outs[1] = {blk(B, m, n, ldb)};
    if (int rc = group_admit(s, ins, outs)) return rc;
    rec_trsm_lln(g_group, m, n, alpha, L, ldl, B, ldb);
    return TTG_CHOL_OK;
  }
  return run_alone(s, [&](Group &g) { rec_trsm_lln(g, m, n, alpha, L, ldl, B, ldb); });
}

int ttg_chol_dtrsm_rln(void *stream, int m, int n, double alpha, const double *L, int ldl, double *B, int ldb) {
  clear_error();
  if (int rc = trsm_args("dtrsm_rln", m, n, n, L, ldl, B, ldb)) return rc;
  if (m == 0 || n == 0) return TTG_CHOL_OK;
  hipStream_t s = as_stream(stream);
  if (grouping(s)) {
    const Span ins[1] = {blk(L, n, n, ldl)}, outs[1] = {blk(B, m, n, ldb)};
    if (int rc = group_admit(s, ins, outs)) return rc;
    rec_trsm_rln(g_group, m, n, alpha, L, ldl, B, ldb);
    return TTG_CHOL_OK;
  }
  return run_alone(s, [&](Group &g) { rec_trsm_rln(g, m, n, alpha, L, ldl, B, ldb); });
}

int ttg_chol_dlauum_tile(void *stream, int n, double *A, int lda) {
  clear_error();
  if (bad_tile(n, lda, A)) {
    set_error("dlauum_tile: bad argument (n=%d lda=%d)", n, lda);
    return TTG_CHOL_EINVAL;
  }
  if (n == 0) return TTG_CHOL_OK;
  hipStream_t s = as_stream(stream);
  if (grouping(s)) {
    const Span none[1], outs[1] = {blk(A, n, n, lda)};
    if (int rc = group_admit(s, none, outs)) return rc;
    rec_lauum(g_group, n, A, lda);
    return TTG_CHOL_OK;
  }
  return run_alone(s, [&](Group &g) { rec_lauum(g, n, A, lda); });
}

int ttg_chol_dtrmm_llt(void *stream, int m, int n, double alpha, const double *L, int ldl, double *B, int ldb) {
  clear_error();
  if (int rc = trsm_args("dtrmm_llt", m, n, m, L, ldl, B, ldb)) return rc;
  if (m == 0 || n == 0) return TTG_CHOL_OK;
  hipStream_t s = as_stream(stream);
  if (grouping(s)) {
    const Span ins[1] = {blk(L, m, m, ldl)}, outs[1] = {blk(B, m, n, ldb)};
    if (int rc = group_admit(s, ins, outs)) return rc;
    rec_trmm_llt(g_group, m, n, alpha, L, ldl, B, ldb);
    return TTG_CHOL_OK;
  }
  return run_alone(s, [&](Group &g) { rec_trmm_llt(g, m, n, alpha, L, ldl, B, ldb); });
}

// C := alpha op(A) op(B) + beta C with op(A) = A (ta = 0, m x k) or A^T
// (ta = 1, A stored k x m) and op(B) = B (k x n); lower: only C's lower
// triangle (C square)
static int update_tb_entry(const char *name, hipStream_t s, bool ta, int m, int n, int k, double alpha,
                           const double *A, int lda, const double *B, int ldb, double beta, double *C, int ldc,
                           int lower) {
  if (m < 0 || n < 0 || k < 0 || lda < std::max(1, ta ? k : m) || ldb < std::max(1, k) || ldc < std::max(1, m) ||
      (m > 0 && n > 0 && (!C || (k > 0 && (!A || !B))))) {
    set_error("%s: bad argument (m=%d n=%d k=%d lda=%d ldb=%d ldc=%d)", name, m, n, k, lda, ldb, ldc);
    return TTG_CHOL_EINVAL;
  }
  if (m == 0 || n == 0) return TTG_CHOL_OK;
  if (k == 0) alpha = 0.0;  // C := beta C
  UpdOp<double> op{A, B, C, m, n, k, lda, ldb, ldc, lower};
  const UpdKey key{alpha, beta, ta, true};
  if (grouping(s)) {
    const Span ins[2] = {ta ? blk(A, k, m, lda) : blk(A, m, k, lda), blk(B, k, n, ldb)},
               outs[1] = {blk(C, m, n, ldc)};
    if (int rc = group_admit(s, ins, outs)) return rc;
    g_group.add_upd(key, op, 0);
    return TTG_CHOL_OK;
  }
  return run_alone(s, [&](Group &g) { g.add_upd(key, op, 0); });
}

int ttg_chol_dgemm_nn(void *stream, int m, int n, int k, double alpha, const double *A, int lda, const double *B,
                      int ldb, double beta, double *C, int ldc) {
  clear_error();
  return update_tb_entry("dgemm_nn", as_stream(stream), false, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, 0);
}

int ttg_chol_dgemm_tn(void *stream, int m, int n, int k, double alpha, const double *A, int lda, const double *B,
                      int ldb, double beta, double *C, int ldc) {
  clear_error();
  return update_tb_entry("dgemm_tn", as_stream(stream), true, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, 0);
}

int ttg_chol_dsyrk_lt(void *stream, int n, int k, double alpha, const double *A, int lda, double beta, double *C,
                      int ldc) {
  clear_error();
  return update_tb_entry("dsyrk_lt", as_stream(stream), true, n, n, k, alpha, A, lda, A, lda, beta, C, ldc, 1);
}

int ttg_chol_group_begin(void *stream) {
  clear_error();
  if (g_group.active) {
    set_error("group_begin: a group is already open on this thread");
    return TTG_CHOL_ESTATE;
  }
  g_group.active = true;
  g_group.stream = as_stream(stream);
  g_group.clear();
  g_group.last_launches = 0;
  g_group.splits = 0;
  return TTG_CHOL_OK;
}

int ttg_chol_group_last_launches(void) { return g_group.last_launches; }

int ttg_chol_group_last_splits(void) { return g_group.splits; }

int ttg_chol_group_end(void *stream) {
  clear_error();
  if (!g_group.active || g_group.stream != as_stream(stream)) {
    set_error("group_end: no group open on this stream");
    return TTG_CHOL_ESTATE;
  }
  g_group.active = false;
  return run_group(g_group, g_group.stream);
}

}  // extern "C"
