// Opaque plan handles: the CSR pattern and the node incidence of a mesh's cells, built once per
// cell content and shared by every caller that asks for the same content -- the C-side form of
// compress_plan's memo (jno/utils/solver/fem_utils.py:4027-4047: a digest of the index array keys
// the cached plan).  The key is (device, the two 64-bit content sums of jno_hash_i32, sizes, nloc,
// vec); a hit is confirmed by comparing the cells with the plan's own copy (jno_equal_i32), so a
// sum collision builds a separate plan instead of returning a wrong one.  Plans are reference
// counted; the last jno_plan_destroy frees the device memory.  A mutex guards the table, and a plan
// is built under it, so two threads asking for one content get one plan.
#include <map>
#include <mutex>
#include <set>
#include <tuple>

#include "common.h"

struct jno_plan {
  int device;
  uint64_t key[2];
  int64_t n_cells, n_nodes, n_rows, nnz;
  int32_t nloc, vec, max_row, refs;
  bool cached;
  int32_t* cells;  // the plan's copy of the cells (the equality check of a hit)
  int32_t* rowptr;
  int32_t* col;
  int32_t* inc_ptr;
  int32_t* inc;
};

namespace jno {

using PlanKey = std::tuple<int, uint64_t, uint64_t, int64_t, int64_t, int32_t, int32_t>;

static std::mutex g_plan_mu;
static std::map<PlanKey, jno_plan*> g_plans;  // cached plans by key
static std::set<jno_plan*> g_live;            // every live handle (cached or not)

static void free_plan(jno_plan* p) {
  if (!p) return;
  (void)hipFree(p->cells);
  (void)hipFree(p->rowptr);
  (void)hipFree(p->col);
  (void)hipFree(p->inc_ptr);
  (void)hipFree(p->inc);
  delete p;
}

// device scratch of one create call, released on every return path
struct PlanScratch {
  void* ws = nullptr;
  void* hash_ws = nullptr;
  ~PlanScratch() {
    (void)hipFree(ws);
    (void)hipFree(hash_ws);
  }
};

static int build_plan(jno_plan* p, const int32_t* d_cells, hipStream_t s, PlanScratch& sc) {
  const int64_t m = p->n_cells * p->nloc;
  JNO_HIP(hipMalloc((void**)&p->cells, sizeof(int32_t) * (size_t)std::max<int64_t>(m, 1)));
  JNO_HIP(hipMalloc((void**)&p->rowptr, sizeof(int32_t) * (size_t)(p->n_rows + 1)));
  JNO_HIP(hipMalloc((void**)&p->inc_ptr, sizeof(int32_t) * (size_t)(p->n_nodes + 1)));
  JNO_HIP(hipMalloc((void**)&p->inc, sizeof(int32_t) * (size_t)std::max<int64_t>(m, 1)));
  if (m > 0) JNO_HIP(hipMemcpyAsync(p->cells, d_cells, sizeof(int32_t) * (size_t)m, hipMemcpyDeviceToDevice, s));
  const size_t wsb = jno_pattern_workspace_size(p->n_cells, p->nloc, p->n_nodes);
  JNO_HIP(hipMalloc(&sc.ws, wsb));
  int64_t nnz = 0;
  int32_t max_row = 0;
  int rc = jno_pattern_count(d_cells, p->n_cells, p->nloc, p->vec, p->n_nodes, p->inc_ptr, p->inc, p->rowptr, &nnz,
                             &max_row, sc.ws, wsb, s);
  if (rc) return rc;
  p->nnz = nnz;
  p->max_row = max_row;
  JNO_HIP(hipMalloc((void**)&p->col, sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1)));
  rc = jno_pattern_fill(d_cells, p->n_cells, p->nloc, p->vec, p->n_nodes, p->inc_ptr, p->inc, p->rowptr, p->col, s);
  if (rc) return rc;
  JNO_HIP(hipStreamSynchronize(s));  // the scratch is freed when the call returns
  return JNO_OK;
}

}  // namespace jno

using namespace jno;

extern "C" int jno_plan_create(const int32_t* d_cells, int64_t n_cells, int nloc, int vec, int64_t n_nodes,
                               jno_plan** h_plan, void* stream) {
  JNO_REQUIRE(h_plan, JNO_EINVAL, "jno_plan_create: null handle pointer");
  *h_plan = nullptr;
  JNO_REQUIRE(d_cells && n_cells > 0 && n_nodes > 0, JNO_EINVAL, "jno_plan_create: empty mesh");
  JNO_REQUIRE(nloc >= 1 && nloc <= 10 && vec >= 1 && vec <= 3, JNO_EINVAL,
              "jno_plan_create: nloc must be in 1..10 and vec in 1..3 (got %d, %d)", nloc, vec);
  JNO_REQUIRE(n_cells * nloc < INT32_MAX && n_nodes * vec < INT32_MAX, JNO_EINVAL, "jno_plan_create: mesh too large");
  hipStream_t s = as_stream(stream);
  int dev = 0;
  JNO_HIP(hipGetDevice(&dev));
  PlanScratch sc;
  JNO_HIP(hipMalloc(&sc.hash_ws, 256));
  uint64_t h[2] = {0, 0};
  int rc = jno_hash_i32(d_cells, n_cells * nloc, h, sc.hash_ws, 256, stream);
  if (rc) return rc;
  const PlanKey key{dev, h[0], h[1], n_cells, n_nodes, nloc, vec};
  std::lock_guard<std::mutex> lock(g_plan_mu);
  auto it = g_plans.find(key);
  bool collision = false;
  if (it != g_plans.end()) {
    int eq = 0;
    rc = jno_equal_i32(d_cells, it->second->cells, n_cells * nloc, &eq, sc.hash_ws, 256, stream);
    if (rc) return rc;
    if (eq) {
      it->second->refs += 1;
      *h_plan = it->second;
      return JNO_OK;
    }
    collision = true;  // equal sums, other content: a plan of its own, not filed under the key
  }
  jno_plan* p = new jno_plan{};
  p->device = dev;
  p->key[0] = h[0];
  p->key[1] = h[1];
  p->n_cells = n_cells;
  p->n_nodes = n_nodes;
  p->n_rows = n_nodes * vec;
  p->nloc = nloc;
  p->vec = vec;
  p->refs = 1;
  p->cached = !collision;
  rc = build_plan(p, d_cells, s, sc);
  if (rc) {
    free_plan(p);
    return rc;
  }
  if (p->cached) g_plans[key] = p;
  g_live.insert(p);
  *h_plan = p;
  return JNO_OK;
}

extern "C" int jno_plan_get(const jno_plan* plan, jno_plan_view* h_view) {
  JNO_REQUIRE(plan && h_view, JNO_EINVAL, "jno_plan_get: null pointer");
  std::lock_guard<std::mutex> lock(g_plan_mu);
  JNO_REQUIRE(g_live.count(const_cast<jno_plan*>(plan)), JNO_EINVAL, "jno_plan_get: not a live plan");
  h_view->n_rows = plan->n_rows;
  h_view->nnz = plan->nnz;
  h_view->n_cells = plan->n_cells;
  h_view->n_nodes = plan->n_nodes;
  h_view->nloc = plan->nloc;
  h_view->vec = plan->vec;
  h_view->max_row = plan->max_row;
  h_view->refs = plan->refs;
  h_view->rowptr = plan->rowptr;
  h_view->col = plan->col;
  h_view->inc_ptr = plan->inc_ptr;
  h_view->inc = plan->inc;
  h_view->key[0] = plan->key[0];
  h_view->key[1] = plan->key[1];
  return JNO_OK;
}

extern "C" int jno_plan_destroy(jno_plan* plan) {
  JNO_REQUIRE(plan, JNO_EINVAL, "jno_plan_destroy: null handle");
  jno_plan* dead = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_plan_mu);
    JNO_REQUIRE(g_live.count(plan), JNO_EINVAL, "jno_plan_destroy: not a live plan");
    if (--plan->refs > 0) return JNO_OK;
    g_live.erase(plan);
    if (plan->cached) {
      const PlanKey key{plan->device, plan->key[0], plan->key[1], plan->n_cells, plan->n_nodes, plan->nloc, plan->vec};
      auto it = g_plans.find(key);
      if (it != g_plans.end() && it->second == plan) g_plans.erase(it);
    }
    dead = plan;
  }
  // the plan's memory is freed on its own device, after the work queued there has finished
  const int pdev = dead->device;
  int cur = 0;
  JNO_HIP(hipGetDevice(&cur));
  if (cur != pdev) JNO_HIP(hipSetDevice(pdev));
  const hipError_t e = hipDeviceSynchronize();
  free_plan(dead);
  if (cur != pdev) (void)hipSetDevice(cur);
  JNO_HIP(e);
  return JNO_OK;
}

extern "C" int jno_plan_live_count(int64_t* h_count) {
  JNO_REQUIRE(h_count, JNO_EINVAL, "jno_plan_live_count: null pointer");
  std::lock_guard<std::mutex> lock(g_plan_mu);
  *h_count = (int64_t)g_live.size();
  return JNO_OK;
}
