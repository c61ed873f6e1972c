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
  bool operator<(const UpdKey &o) const { return std::tie(alpha, beta) < std::tie(o.alpha, o.beta); }
};
struct Group {
  bool active = false;
  hipStream_t stream = nullptr;
  std::vector<PotrfOp> potrf;
  std::vector<TrsmOp<double>> trsm;
  std::vector<PotrfOp> trsm_dinv;  // dinv jobs for the TRSMs (A = L tile)
  std::map<UpdKey, std::vector<UpdOp<double>>> upd;
  struct Norm {
    int n;
    const double *x;
    double *r;
  };
  std::vector<Norm> norm_after;  // norms of memory the group writes
  std::vector<PlgsyOp> plgsy;
  int last_launches = 0;
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
  for (auto &o : g.trsm) written = written || overlaps(x, n, o.B, o.m, o.n, o.ldb);
  for (auto &kv : g.upd)
    for (auto &o : kv.second) written = written || overlaps(x, n, o.C, o.m, o.n, o.ldc);
  if (written) g.norm_after.push_back(Group::Norm{n, x, r});
  return written;
}

bool group_record_plgsy(hipStream_t s, const PlgsyOp &op) {
  if (!grouping(s)) return false;
  g_group.plgsy.push_back(op);
  return true;
}

}  // namespace ttgc

using namespace ttgc;

namespace {
// process-wide schedule knobs; -1: not set (environment or default)
int g_knobs[TTG_CHOL_KNOB_COUNT] = {-1, -1, -1, -1};
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
  static const int lo[TTG_CHOL_KNOB_COUNT] = {2, 0, 1, 0}, hi[TTG_CHOL_KNOB_COUNT] = {8, 64, 8, 1};
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
    g_group.trsm.push_back(op);
    g_group.trsm_dinv.push_back(dj);
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
    g_group.upd[UpdKey{alpha, beta}].push_back(op);
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

int ttg_chol_group_begin(void *stream) {
  clear_error();
  if (g_group.active) {
    set_error("group_begin: a group is already open on this thread");
    return TTG_CHOL_ESTATE;
  }
  g_group.active = true;
  g_group.stream = as_stream(stream);
  g_group.potrf.clear();
  g_group.trsm.clear();
  g_group.trsm_dinv.clear();
  g_group.upd.clear();
  g_group.norm_after.clear();
  g_group.plgsy.clear();
  return TTG_CHOL_OK;
}

int ttg_chol_group_last_launches(void) { return g_group.last_launches; }

int ttg_chol_group_end(void *stream) {
  clear_error();
  if (!g_group.active || g_group.stream != as_stream(stream)) {
    set_error("group_end: no group open on this stream");
    return TTG_CHOL_ESTATE;
  }
  Group &g = g_group;
  g.active = false;
  hipStream_t s = g.stream;
  StreamWork &w = stream_work(s);
  // one staging buffer holds every descriptor array of the group
  size_t off = 0;
  auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
  size_t o_plgsy = off;
  off = align(off + g.plgsy.size() * sizeof(PlgsyOp));
  size_t o_potrf = off;
  off = align(off + g.potrf.size() * sizeof(PotrfOp));
  size_t o_dinvops = off;
  off = align(off + g.trsm_dinv.size() * sizeof(PotrfOp));
  size_t o_trsm = off;
  off = align(off + g.trsm.size() * sizeof(TrsmOp<double>));
  std::vector<size_t> o_upd;
  for (auto &kv : g.upd) {
    o_upd.push_back(off);
    off = align(off + kv.second.size() * sizeof(UpdOp<double>));
  }
  size_t dinv_off = off;
  size_t dinv_bytes = 0;
  for (auto &d : g.trsm_dinv) dinv_bytes += sizeof(double) * dinv_doubles(d.n);
  off = align(off + dinv_bytes);
  if (off == 0) {
    g.last_launches = 0;
    return TTG_CHOL_OK;
  }
  TTGC_HIP(w.ops.reserve(off));
  char *base = static_cast<char *>(w.ops.ptr);
  // inverse-block workspace for the TRSMs lives in the same buffer
  {
    size_t p = dinv_off;
    for (size_t i = 0; i < g.trsm.size(); ++i) {
      g.trsm_dinv[i].dinv = reinterpret_cast<double *>(base + p);
      g.trsm[i].dinv = g.trsm_dinv[i].dinv;
      p += sizeof(double) * dinv_doubles(g.trsm_dinv[i].n);
    }
  }
  TTGC_HIP(w.host_reserve(dinv_off));
  char *host = w.host;
  if (!g.plgsy.empty()) memcpy(host + o_plgsy, g.plgsy.data(), g.plgsy.size() * sizeof(PlgsyOp));
  if (!g.potrf.empty()) memcpy(host + o_potrf, g.potrf.data(), g.potrf.size() * sizeof(PotrfOp));
  if (!g.trsm_dinv.empty())
    memcpy(host + o_dinvops, g.trsm_dinv.data(), g.trsm_dinv.size() * sizeof(PotrfOp));
  if (!g.trsm.empty()) memcpy(host + o_trsm, g.trsm.data(), g.trsm.size() * sizeof(TrsmOp<double>));
  {
    size_t i = 0;
    for (auto &kv : g.upd) {
      memcpy(host + o_upd[i], kv.second.data(), kv.second.size() * sizeof(UpdOp<double>));
      ++i;
    }
  }
  if (dinv_off) {
    TTGC_HIP(hipMemcpyAsync(base, host, dinv_off, hipMemcpyHostToDevice, s));
    TTGC_HIP(hipEventRecord(w.copied, s));
    w.pending = true;
  }
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
  if (!g.trsm.empty()) {
    TTGC_HIP(launch_dinv(s, reinterpret_cast<PotrfOp *>(base + o_dinvops), (int)g.trsm_dinv.size(), none));
    int max_m = 0;
    for (auto &t : g.trsm) max_m = std::max(max_m, t.m);
    TrsmOp<double> tnone{};
    TTGC_HIP(launch_trsm<double>(s, reinterpret_cast<TrsmOp<double> *>(base + o_trsm), (int)g.trsm.size(), tnone,
                                 max_m, nullptr));
    launches += 2;
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
                                     unone, max_m, max_n, kv.first.alpha, kv.first.beta, fast, nullptr));
      ++launches;
      ++i;
    }
  }
  for (auto &nm : g.norm_after) TTGC_HIP(launch_nrm2(s, nm.n, nm.x, nm.r));
  g.last_launches = launches;
  return TTG_CHOL_OK;
}

}  // extern "C"
