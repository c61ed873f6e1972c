#include "chol_profile.hpp"

#include <mutex>
#include <vector>

#include "chol_internal.hpp"

namespace ttgc {

struct ProfRec {
  hipEvent_t a = nullptr, b = nullptr;
  int kind = 0;
  double flops = 0.0;
  bool used = false;
};
static std::mutex g_prof_mu;
static std::vector<ProfRec> g_prof;
static size_t g_prof_n = 0;
static bool g_prof_on = false;

bool prof_enabled() { return g_prof_on; }

int prof_begin(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (g_prof_n == g_prof.size()) {
    ProfRec r;
    if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return -1;
    g_prof.push_back(r);
  }
  const int slot = (int)g_prof_n++;
  g_prof[slot].used = false;
  (void)hipEventRecord(g_prof[slot].a, s);
  return slot;
}

void prof_end(int slot, hipStream_t s, int kind, double flops) {
  if (slot < 0) return;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  (void)hipEventRecord(g_prof[slot].b, s);
  g_prof[slot].kind = kind;
  g_prof[slot].flops = flops;
  g_prof[slot].used = true;
}

}  // namespace ttgc

using namespace ttgc;

extern "C" int ttg_chol_profile_enable(int on) {
  g_prof_on = on != 0;
  return TTG_CHOL_OK;
}

extern "C" int ttg_chol_profile_read(int kind, double *ms_total, long long *launches, double *flops) {
  clear_error();
  if (kind < 0 || kind >= PROF_KINDS || !ms_total || !launches || !flops) {
    set_error("profile_read: bad argument (kind=%d)", kind);
    return TTG_CHOL_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_prof_mu);
  double ms = 0.0, fl = 0.0;
  long long n = 0;
  for (size_t i = 0; i < g_prof_n; ++i) {
    ProfRec &r = g_prof[i];
    if (!r.used || r.kind != kind) continue;
    TTGC_HIP(hipEventSynchronize(r.b));
    float t = 0.f;
    TTGC_HIP(hipEventElapsedTime(&t, r.a, r.b));
    ms += t;
    fl += r.flops;
    ++n;
    r.used = false;
  }
  *ms_total = ms;
  *launches = n;
  *flops = fl;
  bool any = false;
  for (size_t i = 0; i < g_prof_n; ++i) any = any || g_prof[i].used;
  if (!any) g_prof_n = 0;  // every slot consumed: recycle the events
  return TTG_CHOL_OK;
}
