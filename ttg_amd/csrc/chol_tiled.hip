// extern "C" FP64 tiled driver over TiledPlanT<double> (the examples/potrf
// task graph with per-step launch grouping, lookahead and hipGraph replay).
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "chol_tiled_plan.hpp"

namespace ttgc {

static std::mutex g_plan_mu;
static std::map<std::tuple<int, int, double *, unsigned>, std::unique_ptr<TiledPlanT<double>>> g_plans;

void release_tiled_plans() {
  std::lock_guard<std::mutex> lk(g_plan_mu);
  g_plans.clear();
}

}  // namespace ttgc

using namespace ttgc;

extern "C" int ttg_chol_dpotrf_tiled(void *stream, int N, int NB, double *tiles, int *d_info, unsigned flags) {
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
    plan.reset(new TiledPlanT<double>);
    plan->N = N;
    plan->NB = NB;
    plan->tiles = tiles;
    plan->flags = flags;
    int rc = plan->build();
    if (rc != TTG_CHOL_OK) {
      g_plans.erase(key);
      return rc;
    }
  }
  int *info = d_info ? d_info : static_cast<int *>(plan->d_info.ptr);
  return plan->run(as_stream(stream), info);
}
