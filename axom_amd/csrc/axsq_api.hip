// axsq_api.hip -- the extern "C" entry points declared in include/axsq.h.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include <algorithm>
#include <cstddef>
#include <cmath>

#include "axsq_internal.hpp"

namespace axsq {

int bvh_export_launch(const axsq_bvh* b, double* inner_nodes, int32_t* children,
                      int32_t* leaf_nodes, uint32_t* mcodes, hipStream_t s);
int sdf_check_conn(const axsq_sdf* h, hipStream_t s);
int sdf_setup(axsq_sdf* h, hipStream_t s);
int sdf_build_v2t(axsq_sdf* h, hipStream_t s);
int sdf_eval_launch(const axsq_sdf* h, const double* x, const double* y, const double* z,
                    int64_t stride, int64_t nq, int fp_mode, double* dist, double* cp,
                    double* normal, int32_t* elem, bool stats, hipStream_t s);
int sdf_backward_launch(const axsq_sdf* h, const double* x, const double* y, const double* z,
                        int64_t stride, int64_t nq, const double* dist, const int32_t* elem,
                        const float* gout, double* grad_q, double* grad_v, hipStream_t s);

// ---- errors -------------------------------------------------------------------
static thread_local std::string t_err;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  t_err = buf;
}
void clear_error() { t_err.clear(); }

// ---- device buffers / workspace ---------------------------------------------
int DevBuf::ensure(size_t b) {
  if (b <= bytes && p) return AXSQ_OK;
  if (p) hipFree(p);
  p = nullptr;
  bytes = 0;
  size_t want = b + b / 4 + 4096;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    p = nullptr;
    set_error("workspace hipMalloc of %zu bytes failed: %s", want, hipGetErrorString(e));
    return AXSQ_ERR_NOMEM;
  }
  bytes = want;
  return AXSQ_OK;
}
void DevBuf::release() {
  if (p) hipFree(p);
  p = nullptr;
  bytes = 0;
}
void Workspace::release() {
  k0.release();
  v0.release();
  k1.release();
  v1.release();
  hist.release();
  tiles.release();
  misc.release();
  flag.release();
  redo.release();
  susp.release();
  cnt.release();
  mesh_tri.release();
  mesh_box.release();
  mesh_int.release();
  mesh_slots.release();
  hot.release();
  wave_cost.release();
  wave_order.release();
  find_stash.release();
  find_cnt.release();
  find_over.release();
  find_n.release();
  one.release();
  if (mesh_bvh) {
    axsq_bvh_destroy(mesh_bvh);
    mesh_bvh = nullptr;
  }
}

namespace {
struct ThreadWorkspaces {
  std::vector<Workspace> per_device;
  // Device memory is left to the runtime at thread exit: freeing here could
  // run after the HIP runtime has been torn down.
};
thread_local ThreadWorkspaces t_ws;
}  // namespace

Workspace& workspace() {
  int dev = 0;
  hipGetDevice(&dev);
  if ((int)t_ws.per_device.size() <= dev) t_ws.per_device.resize(dev + 1);
  Workspace& w = t_ws.per_device[dev];
  w.device = dev;
  return w;
}

// ---- profiling -------------------------------------------------------------
namespace {
std::mutex g_prof_mu;
bool g_prof_on = false;
struct Pending {
  hipEvent_t e0, e1;
  std::string name;
};
std::vector<Pending> g_pending;
std::map<std::string, std::pair<int64_t, double>> g_totals;
}  // namespace

bool prof_on() { return g_prof_on; }

ProfRange::ProfRange(const char* n, hipStream_t st) : name(n), s(st) {
  if (!g_prof_on) return;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    e0 = e1 = nullptr;
    return;
  }
  hipEventRecord(e0, s);
}

void ProfRange::stop() {
  if (!e0) return;
  hipEventRecord(e1, s);
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_pending.push_back({e0, e1, name});
  e0 = e1 = nullptr;
}

void prof_flush() {
  std::lock_guard<std::mutex> g(g_prof_mu);
  std::vector<Pending> keep;
  for (auto& p : g_pending) {
    if (hipEventQuery(p.e1) != hipSuccess) {
      keep.push_back(p);
      continue;
    }
    float ms = 0.f;
    hipEventElapsedTime(&ms, p.e0, p.e1);
    auto& t = g_totals[p.name];
    t.first += 1;
    t.second += ms;
    hipEventDestroy(p.e0);
    hipEventDestroy(p.e1);
  }
  g_pending.swap(keep);
}

// Waits for the stream, folds profiling ranges, and turns a traversal stack
// overflow flagged by a kernel into an error code.
static int finish(int32_t* status, hipStream_t s) {
  AXSQ_CHECK_HIP(hipStreamSynchronize(s));
  prof_flush();
  if (status) {
    int32_t h = 0;
    AXSQ_CHECK_HIP(hipMemcpy(&h, status, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (h) {
      int32_t z = 0;
      AXSQ_CHECK_HIP(hipMemcpy(status, &z, sizeof(int32_t), hipMemcpyHostToDevice));
      if (h & kStatusStack) {
        set_error("traversal stopped: stack deeper than %d entries or malformed tree", kStack);
        return AXSQ_ERR_STACK_OVERFLOW;
      }
      if (h & kStatusFillMismatch) {
        set_error("candidate fill: the search found another number of candidates than the "
                  "counts give for some query (queries, tolerance or tree changed since the "
                  "count); no id was written past a query's count");
        return AXSQ_ERR_INVALID_ARG;
      }
      set_error("element id out of range in the adjoint input");
      return AXSQ_ERR_INVALID_ARG;
    }
  }
  return AXSQ_OK;
}

static bool valid_kind(int k) {
  return k == AXSQ_QUERY_POINTS || k == AXSQ_QUERY_BOXES || k == AXSQ_QUERY_RAYS;
}

// f32q: the queries are floats (the _f32 entry points), which a float tree
// takes and a double tree refuses, and the other way round.
static int check_find_args(const axsq_bvh* b, int kind, const void* x, const void* y,
                           const void* z, int64_t stride, int64_t nq, bool f32q) {
  if (!b || !b->nodes) {
    set_error("BVH handle is NULL or not built");
    return AXSQ_ERR_STATE;
  }
  if (b->f32 != f32q) {
    set_error(b->f32 ? "a float tree takes float queries (the axsq_bvh_find*_f32 entry points)"
                     : "a double tree takes double queries (the axsq_bvh_find* entry points "
                       "without _f32)");
    return AXSQ_ERR_STATE;
  }
  if (!valid_kind(kind)) {
    set_error("unknown query kind %d", kind);
    return AXSQ_ERR_INVALID_ARG;
  }
  if (nq < 0 || nq > 0x7fffffffll) {
    set_error("query count %lld out of range", (long long)nq);
    return AXSQ_ERR_INVALID_ARG;
  }
  if (nq > 0 && !x) {
    set_error("query pointer is NULL");
    return AXSQ_ERR_INVALID_ARG;
  }
  if (kind == AXSQ_QUERY_POINTS) {
    if (nq > 0 && (!y || (!z && b->dims == 3))) {
      set_error("point queries need x, y (and z for a 3-D tree) pointers");
      return AXSQ_ERR_INVALID_ARG;
    }
    if (stride < 1) {
      set_error("point stride must be >= 1");
      return AXSQ_ERR_INVALID_ARG;
    }
  } else if (stride < 2 * b->dims) {
    set_error("box/ray queries need a stride of at least %d doubles", 2 * b->dims);
    return AXSQ_ERR_INVALID_ARG;
  }
  return AXSQ_OK;
}

}  // namespace axsq

using namespace axsq;

extern "C" {

const char* axsq_last_error(void) { return t_err.c_str(); }

int axsq_version(void) { return 100; }

int axsq_free(void* p) {
  if (!p) return AXSQ_OK;
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) == hipSuccess && attr.type == hipMemoryTypeHost) {
    AXSQ_CHECK_HIP(hipHostFree(p));  // a candidate array in pinned host memory
    return AXSQ_OK;
  }
  (void)hipGetLastError();
  AXSQ_CHECK_HIP(hipFree(p));
  return AXSQ_OK;
}

// ---- BVH ------------------------------------------------------------------
int axsq_bvh_build(axsq_bvh** out, const double* boxes, int64_t n, double scale, int fp_mode,
                   void* stream) {
  clear_error();
  if (!out) {
    set_error("axsq_bvh_build: out is NULL");
    return AXSQ_ERR_INVALID_ARG;
  }
  *out = nullptr;
  if (fp_mode != AXSQ_FP64 && fp_mode != AXSQ_FP32_TRAVERSAL) {
    set_error("axsq_bvh_build: unknown fp_mode %d", fp_mode);
    return AXSQ_ERR_INVALID_ARG;
  }
  axsq_bvh* b = new axsq_bvh();
  hipGetDevice(&b->device);
  const int rc = bvh_build_into(b, boxes, n, scale, fp_mode, (hipStream_t)stream);
  if (rc != AXSQ_OK) {
    bvh_release(b);
    delete b;
    return rc;
  }
  *out = b;
  return AXSQ_OK;
}

int axsq_bvh_build2d(axsq_bvh** out, const double* boxes, int64_t n, double scale, int fp_mode,
                     void* stream) {
  clear_error();
  if (!out) {
    set_error("axsq_bvh_build2d: out is NULL");
    return AXSQ_ERR_INVALID_ARG;
  }
  *out = nullptr;
  if (fp_mode != AXSQ_FP64 && fp_mode != AXSQ_FP32_TRAVERSAL) {
    set_error("axsq_bvh_build2d: unknown fp_mode %d", fp_mode);
    return AXSQ_ERR_INVALID_ARG;
  }
  if (n < 0 || (n > 0 && !boxes)) {
    set_error("axsq_bvh_build2d: need n >= 0 boxes");
    return AXSQ_ERR_INVALID_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  Workspace& ws = workspace();
  AXSQ_TRY(ws.box2d.ensure(sizeof(double) * 6 * (size_t)(n > 0 ? n : 1)));
  AXSQ_TRY(widen_boxes2d(boxes, n, ws.box2d.as<double>(), s));
  axsq_bvh* b = new axsq_bvh();
  hipGetDevice(&b->device);
  const int rc = bvh_build_into(b, ws.box2d.as<double>(), n, scale, fp_mode, s, 2);
  if (rc != AXSQ_OK) {
    bvh_release(b);
    delete b;
    return rc;
  }
  *out = b;
  return AXSQ_OK;
}

int axsq_bvh_rebuild(axsq_bvh* b, const double* boxes, int64_t n, double scale, int fp_mode,
                     void* stream) {
  clear_error();
  if (!b) {
    set_error("axsq_bvh_rebuild: bvh is NULL");
    return AXSQ_ERR_INVALID_ARG;
  }
  if (fp_mode != AXSQ_FP64 && fp_mode != AXSQ_FP32_TRAVERSAL) {
    set_error("axsq_bvh_rebuild: unknown fp_mode %d", fp_mode);
    return AXSQ_ERR_INVALID_ARG;
  }
  if (n < 0 || (n > 0 && !boxes)) {
    set_error("axsq_bvh_rebuild: need n >= 0 boxes");
    return AXSQ_ERR_INVALID_ARG;
  }
  int dev = -1;
  AXSQ_CHECK_HIP(hipGetDevice(&dev));
  if (dev != b->device) {
    set_error("axsq_bvh_rebuild: the tree lives on device %d, the current device is %d",
              b->device, dev);
    return AXSQ_ERR_INVALID_ARG;
  }
  if (b->f32) {
    set_error("axsq_bvh_rebuild: a float tree is rebuilt with axsq_bvh_rebuild_f32");
    return AXSQ_ERR_STATE;
  }
  hipStream_t s = (hipStream_t)stream;
  if (b->dims == 2) {
    Workspace& ws = workspace();
    AXSQ_TRY(ws.box2d.ensure(sizeof(double) * 6 * (size_t)(n > 0 ? n : 1)));
    AXSQ_TRY(widen_boxes2d(boxes, n, ws.box2d.as<double>(), s));
    return bvh_build_into(b, ws.box2d.as<double>(), n, scale, fp_mode, s, 2);
  }
  return bvh_build_into(b, boxes, n, scale, fp_mode, s, 3);
}

int axsq_bvh_build_f32(axsq_bvh** out, const float* boxes, int64_t n, int ndims, float scale,
                       void* stream) {
  clear_error();
  if (!out) {
    set_error("axsq_bvh_build_f32: out is NULL");
    return AXSQ_ERR_INVALID_ARG;
  }
  *out = nullptr;
  if (ndims != 2 && ndims != 3) {
    set_error("axsq_bvh_build_f32: ndims must be 2 or 3, not %d", ndims);
    return AXSQ_ERR_INVALID_ARG;
  }
  axsq_bvh* b = new axsq_bvh();
  hipGetDevice(&b->device);
  const int rc = bvh_build_into(b, nullptr, n, (double)scale, AXSQ_FP32_TRAVERSAL,
                                (hipStream_t)stream, ndims, true, boxes);
  if (rc != AXSQ_OK) {
    bvh_release(b);
    delete b;
    return rc;
  }
  *out = b;
  return AXSQ_OK;
}

int axsq_bvh_rebuild_f32(axsq_bvh* b, const float* boxes, int64_t n, float scale,
                         void* stream) {
  clear_error();
  if (!b) {
    set_error("axsq_bvh_rebuild_f32: bvh is NULL");
    return AXSQ_ERR_INVALID_ARG;
  }
  if (!b->f32) {
    set_error("axsq_bvh_rebuild_f32: a double tree is rebuilt with axsq_bvh_rebuild");
    return AXSQ_ERR_STATE;
  }
  int dev = -1;
  AXSQ_CHECK_HIP(hipGetDevice(&dev));
  if (dev != b->device) {
    set_error("axsq_bvh_rebuild_f32: the tree lives on device %d, the current device is %d",
              b->device, dev);
    return AXSQ_ERR_INVALID_ARG;
  }
  return bvh_build_into(b, nullptr, n, (double)scale, AXSQ_FP32_TRAVERSAL, (hipStream_t)stream,
                        b->dims, true, boxes);
}

int axsq_bvh_precision(const axsq_bvh* b) { return b ? (b->f32 ? 32 : 64) : 0; }

int axsq_bvh_dimension(const axsq_bvh* b) { return b ? b->dims : 0; }

int axsq_bvh_set_allocator(axsq_bvh* b, int allocator_id) {
  clear_error();
  if (!b) {
    set_error("axsq_bvh_set_allocator: bvh is NULL");
    return AXSQ_ERR_INVALID_ARG;
  }
  if (allocator_id != AXSQ_ALLOC_DEVICE && allocator_id != AXSQ_ALLOC_MANAGED &&
      allocator_id != AXSQ_ALLOC_PINNED_HOST) {
    set_error("axsq_bvh_set_allocator: unknown allocator id %d", allocator_id);
    return AXSQ_ERR_INVALID_ARG;
  }
  b->allocator = allocator_id;
  return AXSQ_OK;
}

int axsq_bvh_get_allocator(const axsq_bvh* b) { return b ? b->allocator : AXSQ_ALLOC_DEVICE; }

int axsq_release_workspace(void) {
  clear_error();
  workspace().release();
  return AXSQ_OK;
}

static_assert(sizeof(axsq_node) == sizeof(axsq::Node64), "axsq_node mirrors Node64");
static_assert(offsetof(axsq_node, lc) == offsetof(axsq::Node64, lc) &&
                  offsetof(axsq_node, flags) == offsetof(axsq::Node64, flags) &&
                  offsetof(axsq_node, lo) == offsetof(axsq::Node64, lo) &&
                  offsetof(axsq_node, mid) == offsetof(axsq::Node64, mid),
              "axsq_node mirrors Node64");

int axsq_bvh_device_arrays(const axsq_bvh* b, const axsq_node** nodes, const int32_t** leaf_ids,
                           int32_t* ninner) {
  clear_error();
  if (!b || !nodes || !leaf_ids || !ninner) {
    set_error("axsq_bvh_device_arrays: NULL argument");
    return AXSQ_ERR_INVALID_ARG;
  }
  if (!b->nodes) {
    set_error("axsq_bvh_device_arrays: the tree is not built");
    return AXSQ_ERR_STATE;
  }
  *nodes = reinterpret_cast<const axsq_node*>(b->nodes);
  *leaf_ids = b->leaf_ids;
  *ninner = (int32_t)(b->n - 1);
  return AXSQ_OK;
}

int axsq_bvh_destroy(axsq_bvh* b) {
  if (!b) return AXSQ_OK;
  bvh_release(b);
  delete b;
  return AXSQ_OK;
}

int axsq_bvh_bounds(const axsq_bvh* b, double lo[3], double hi[3]) {
  clear_error();
  if (!b || !lo || !hi) {
    set_error("axsq_bvh_bounds: NULL argument");
    return AXSQ_ERR_INVALID_ARG;
  }
  for (int d = 0; d < 3; ++d) {
    lo[d] = b->bounds[d];
    hi[d] = b->bounds[3 + d];
  }
  return AXSQ_OK;
}

int64_t axsq_bvh_num_leaves(const axsq_bvh* b) { return b ? b->n : 0; }

int axsq_bvh_export(const axsq_bvh* b, double* inner_nodes, int32_t* children,
                    int32_t* leaf_nodes, uint32_t* mcodes, void* stream) {
  clear_error();
  if (!b || !b->nodes) {
    set_error("axsq_bvh_export: BVH not built");
    return AXSQ_ERR_STATE;
  }
  hipStream_t s = (hipStream_t)stream;
  AXSQ_TRY(bvh_export_launch(b, inner_nodes, children, leaf_nodes, mcodes, s));
  return finish(nullptr, s);
}

// The count pass of a candidate search: counts, offsets and *total. With
// keep, also the first stash_slots(nq) ids of every query and the list of
// queries with more, for the fill of the same call (axsq_bvh_find).
static int find_count_impl(const axsq_bvh* b, int kind, const void* x, const void* y,
                           const void* z, int64_t stride, int64_t nq, double tol,
                           int32_t* offsets, int32_t* counts, int64_t* total, bool keep,
                           int64_t* nover_out, int* kstash_out, hipStream_t s) {
  *total = 0;
  if (nq == 0) return AXSQ_OK;
  Workspace& ws = workspace();
  const int kstash = keep ? stash_slots(nq) : 0;
  if (keep) {
    AXSQ_TRY(ws.find_stash.ensure(sizeof(int32_t) * kstash * (size_t)nq));
    AXSQ_TRY(ws.find_cnt.ensure(sizeof(int32_t) * (size_t)nq));
    AXSQ_TRY(ws.find_over.ensure(sizeof(int32_t) * (size_t)nq));
  }
  AXSQ_TRY(ws.find_n.ensure(sizeof(int64_t) * 2));
  // find_n: [0] the scanned total (int64), [1] the listed queries (uint32)
  int64_t* tot_dev = ws.find_n.as<int64_t>();
  uint32_t* nover = reinterpret_cast<uint32_t*>(tot_dev + 1);
  AXSQ_TRY(fill_i32(reinterpret_cast<int32_t*>(nover), 0, 1, s));
  if (b->stats) AXSQ_TRY(fill_i32(reinterpret_cast<int32_t*>(b->find_visits), 0, 2, s));
  AXSQ_TRY(find_launch(b, kind, 0, x, y, z, stride, nq, tol, counts, nullptr, nullptr, nullptr,
                       keep ? ws.find_stash.as<int32_t>() : nullptr, kstash,
                       keep ? ws.find_cnt.as<int32_t>() : nullptr,
                       keep ? ws.find_over.as<int32_t>() : nullptr, nover, s));
  {
    ProfRange ps("bvh_find_scan", s);
    AXSQ_TRY(scan_exclusive_i32(counts, offsets, nq, tot_dev, ws, s));
  }
  int64_t tn[2] = {0, 0};
  AXSQ_CHECK_HIP(hipMemcpyAsync(tn, tot_dev, sizeof(tn), hipMemcpyDeviceToHost, s));
  AXSQ_TRY(finish(b->status, s));
  if (b->stats) {
    unsigned long long v = 0;
    AXSQ_CHECK_HIP(hipMemcpy(&v, b->find_visits, sizeof(v), hipMemcpyDeviceToHost));
    const_cast<axsq_bvh*>(b)->last_visits = (int64_t)v;
  }
  const int64_t t = tn[0];
  if (t > 0x7fffffffll) {
    set_error("candidate total %lld exceeds the int32 index range", (long long)t);
    return AXSQ_ERR_INVALID_ARG;
  }
  *total = t;
  if (nover_out) *nover_out = (int64_t)(uint32_t)tn[1];
  if (kstash_out) *kstash_out = kstash;
  return AXSQ_OK;
}

}  // extern "C"

static int find_count_any(const axsq_bvh* b, int kind, const void* x, const void* y,
                          const void* z, int64_t stride, int64_t nq, double tol,
                          int32_t* offsets, int32_t* counts, int64_t* total, void* stream,
                          bool f32q) {
  clear_error();
  AXSQ_TRY(check_find_args(b, kind, x, y, z, stride, nq, f32q));
  if (!total || (nq > 0 && (!offsets || !counts))) {
    set_error("axsq_bvh_find_count: NULL output");
    return AXSQ_ERR_INVALID_ARG;
  }
  return find_count_impl(b, kind, x, y, z, stride, nq, tol, offsets, counts, total, false,
                         nullptr, nullptr, (hipStream_t)stream);
}

static int find_fill_any(const axsq_bvh* b, int kind, const void* x, const void* y,
                         const void* z, int64_t stride, int64_t nq, double tol,
                         const int32_t* offsets, const int32_t* counts, int32_t* candidates,
                         void* stream, bool f32q) {
  clear_error();
  AXSQ_TRY(check_find_args(b, kind, x, y, z, stride, nq, f32q));
  if (nq > 0 && (!offsets || !counts)) {
    set_error("axsq_bvh_find_fill: offsets or counts is NULL");
    return AXSQ_ERR_INVALID_ARG;
  }
  if (nq == 0) return AXSQ_OK;
  hipStream_t s = (hipStream_t)stream;
  AXSQ_TRY(find_launch(b, kind, 1, x, y, z, stride, nq, tol, nullptr, counts, offsets,
                       candidates, nullptr, 0, nullptr, nullptr, nullptr, s));
  return finish(b->status, s);
}

// Candidate array of *total ids (at least one) with the allocator the tree
// was given (axsq_bvh_set_allocator).
static int alloc_candidates(const axsq_bvh* b, int64_t total, int32_t** out) {
  const size_t bytes = sizeof(int32_t) * (size_t)(total > 0 ? total : 1);
  hipError_t e;
  if (b->allocator == AXSQ_ALLOC_MANAGED)
    e = hipMallocManaged((void**)out, bytes);
  else if (b->allocator == AXSQ_ALLOC_PINNED_HOST)
    e = hipHostMalloc((void**)out, bytes, hipHostMallocDefault);
  else
    e = hipMalloc((void**)out, bytes);
  if (e != hipSuccess) {
    *out = nullptr;
    set_error("candidate allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    return AXSQ_ERR_NOMEM;
  }
  return AXSQ_OK;
}

// The fill of a one-call search: the ids its count pass kept, then a search
// for the nover listed queries with more.
static int find_fill_kept(const axsq_bvh* b, int kind, const void* x, const void* y,
                          const void* z, int64_t stride, int64_t nq, double tol,
                          const int32_t* offsets, const int32_t* counts, int32_t* cands,
                          int64_t nover, int kstash, hipStream_t s) {
  Workspace& ws = workspace();
  AXSQ_TRY(stash_copy_launch(ws.find_stash.as<int32_t>(), kstash, ws.find_cnt.as<int32_t>(),
                             offsets, nq, cands, s));
  if (nover > 0)
    AXSQ_TRY(find_launch(b, kind, 2, x, y, z, stride, nover, tol, nullptr, counts, offsets,
                         cands, nullptr, 0, nullptr, ws.find_over.as<int32_t>(), nullptr, s));
  return finish(b->status, s);
}

static int check_find_outputs(const char* name, int64_t nq, const int32_t* offsets,
                              const int32_t* counts, const int64_t* total) {
  if (!total || (nq > 0 && (!offsets || !counts))) {
    set_error("%s: NULL offsets, counts or total", name);
    return AXSQ_ERR_INVALID_ARG;
  }
  return AXSQ_OK;
}

static int find_into_any(const axsq_bvh* b, int kind, const void* x, const void* y,
                         const void* z, int64_t stride, int64_t nq, double tol, int32_t* offsets,
                         int32_t* counts, int32_t* candidates, int64_t capacity, int64_t* total,
                         void* stream, bool f32q) {
  clear_error();
  AXSQ_TRY(check_find_args(b, kind, x, y, z, stride, nq, f32q));
  AXSQ_TRY(check_find_outputs("axsq_bvh_find_into", nq, offsets, counts, total));
  *total = 0;
  if (capacity < 0 || (capacity > 0 && !candidates)) {
    set_error("axsq_bvh_find_into: negative capacity or NULL candidates");
    return AXSQ_ERR_INVALID_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  int64_t nover = 0;
  int kstash = 0;
  AXSQ_TRY(find_count_impl(b, kind, x, y, z, stride, nq, tol, offsets, counts, total, true,
                           &nover, &kstash, s));
  if (*total == 0 || *total > capacity) return AXSQ_OK;  // too small: nothing written
  return find_fill_kept(b, kind, x, y, z, stride, nq, tol, offsets, counts, candidates, nover,
                        kstash, s);
}

// ---- the search into caller storage without allocation ----------------------
// axsq_bvh_find_into_noalloc keeps its scratch (the count pass's kept ids,
// their counts, the list of queries with more, the scan's tile sums and the
// totals) in this array, one copy per GPU, held until the process ends.
// Layout, for nq queries: [0, 16) total (int64) and listed-query count
// (uint32); [kNoAllocHeader, +8 * scan_tiles(nq)) tile sums; then nq kept
// counts, nq listed queries and kstash * nq kept ids (int32).
namespace {
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr size_t kNoAllocHeader = 4096;
constexpr int64_t kNoAllocMaxQueries = int64_t(1) << 20;
constexpr int kNoAllocMaxDevices = 64;
static_assert(kNoAllocHeader >= 16 + 8 * ((kNoAllocMaxQueries + kTile - 1) / kTile),
              "tile sums of the largest search fit the header");
static_assert(kNoAllocHeader + 12 * size_t(kNoAllocMaxQueries) <= kNoAllocScratchBytes,
              "the largest search keeps at least one id per query");

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kNoAllocMaxDevices] = {};

// The device address of this device's copy of `scratch`, looked up on the
// device's first call.
int scratch_on_current_device(unsigned char** out) {
  int dev = 0;
  AXSQ_CHECK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) {
    set_error("axsq_bvh_find_into_noalloc: device %d beyond the %d it keeps scratch for", dev,
              kNoAllocMaxDevices);
    return AXSQ_ERR_INVALID_ARG;
  }
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("axsq_bvh_find_into_noalloc: hipGetSymbolAddress(scratch) failed: %s",
                hipGetErrorString(e));
      return AXSQ_ERR_HIP;
    }
    g_scratch_dev[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_dev[dev];
  return AXSQ_OK;
}
}  // namespace

extern "C" {

int axsq_bvh_find_into_noalloc(const axsq_bvh* b, int kind, const double* x, const double* y,
                               const double* z, int64_t stride, int64_t nq, double tol,
                               int32_t* offsets, int32_t* counts, int32_t* candidates,
                               int64_t capacity, int64_t* total, void* stream) {
  clear_error();
  AXSQ_TRY(check_find_args(b, kind, x, y, z, stride, nq, false));
  AXSQ_TRY(check_find_outputs("axsq_bvh_find_into_noalloc", nq, offsets, counts, total));
  *total = 0;
  if (nq > kNoAllocMaxQueries) {
    set_error("axsq_bvh_find_into_noalloc: %lld queries exceed its maximum of %lld (use "
              "axsq_bvh_find_into, or split the queries)",
              (long long)nq, (long long)kNoAllocMaxQueries);
    return AXSQ_ERR_INVALID_ARG;
  }
  if (capacity < 0 || (capacity > 0 && !candidates)) {
    set_error("axsq_bvh_find_into_noalloc: negative capacity or NULL candidates");
    return AXSQ_ERR_INVALID_ARG;
  }
  if (nq == 0) return AXSQ_OK;
  unsigned char* base = nullptr;
  AXSQ_TRY(scratch_on_current_device(&base));
  hipStream_t s = (hipStream_t)stream;
  int64_t* tot_dev = reinterpret_cast<int64_t*>(base);
  uint32_t* nover_dev = reinterpret_cast<uint32_t*>(base + 8);
  int64_t* sums = reinterpret_cast<int64_t*>(base + 16);
  int32_t* kept_cnt = reinterpret_cast<int32_t*>(base + kNoAllocHeader);
  int32_t* over = kept_cnt + nq;
  int32_t* kept = over + nq;
  const size_t room = kNoAllocScratchBytes - kNoAllocHeader - 8 * (size_t)nq;
  const int kstash = (int)std::min<size_t>(64, room / (4 * (size_t)nq));  // >= 1
  // count pass (keeping the first kstash ids), scan, then the fill from the
  // kept ids and a search again for the listed queries
  AXSQ_TRY(fill_i32(reinterpret_cast<int32_t*>(nover_dev), 0, 1, s));
  AXSQ_TRY(find_launch(b, kind, 0, x, y, z, stride, nq, tol, counts, nullptr, nullptr, nullptr,
                       kept, kstash, kept_cnt, over, nover_dev, s));
  AXSQ_TRY(scan_exclusive_i32_into(counts, offsets, nq, sums, tot_dev, s));
  int64_t tn[2] = {0, 0};
  AXSQ_CHECK_HIP(hipMemcpyAsync(tn, tot_dev, sizeof(tn), hipMemcpyDeviceToHost, s));
  AXSQ_TRY(finish(b->status, s));
  if (tn[0] > 0x7fffffffll) {
    set_error("candidate total %lld exceeds the int32 index range", (long long)tn[0]);
    return AXSQ_ERR_INVALID_ARG;
  }
  *total = tn[0];
  if (*total == 0 || *total > capacity) return AXSQ_OK;  // too small: nothing written
  AXSQ_TRY(stash_copy_launch(kept, kstash, kept_cnt, offsets, nq, candidates, s));
  const int64_t nover = (int64_t)(uint32_t)tn[1];
  if (nover > 0)
    AXSQ_TRY(find_launch(b, kind, 2, x, y, z, stride, nover, tol, nullptr, counts, offsets,
                         candidates, nullptr, 0, nullptr, over, nullptr, s));
  return finish(b->status, s);
}

}  // extern "C"

static int find_alloc_any(const axsq_bvh* b, int kind, const void* x, const void* y,
                          const void* z, int64_t stride, int64_t nq, double tol, int32_t* offsets,
                          int32_t* counts, int32_t** candidates, int64_t* total, void* stream,
                          bool f32q) {
  clear_error();
  if (!candidates || !total) {
    set_error("axsq_bvh_find: NULL output");
    return AXSQ_ERR_INVALID_ARG;
  }
  *candidates = nullptr;
  *total = 0;
  AXSQ_TRY(check_find_args(b, kind, x, y, z, stride, nq, f32q));
  AXSQ_TRY(check_find_outputs("axsq_bvh_find", nq, offsets, counts, total));
  hipStream_t s = (hipStream_t)stream;
  int64_t nover = 0;
  int kstash = 0;
  AXSQ_TRY(find_count_impl(b, kind, x, y, z, stride, nq, tol, offsets, counts, total, true,
                           &nover, &kstash, s));
  int32_t* c = nullptr;
  AXSQ_TRY(alloc_candidates(b, *total, &c));
  if (*total > 0) {
    const int rc = find_fill_kept(b, kind, x, y, z, stride, nq, tol, offsets, counts, c, nover,
                                  kstash, s);
    if (rc != AXSQ_OK) {
      (void)axsq_free(c);
      *total = 0;
      return rc;
    }
  }
  *candidates = c;
  return AXSQ_OK;
}

extern "C" {

int axsq_bvh_find_count(const axsq_bvh* b, int kind, const double* x, const double* y,
                        const double* z, int64_t stride, int64_t nq, double tol,
                        int32_t* offsets, int32_t* counts, int64_t* total, void* stream) {
  return find_count_any(b, kind, x, y, z, stride, nq, tol, offsets, counts, total, stream, false);
}
int axsq_bvh_find_count_f32(const axsq_bvh* b, int kind, const float* x, const float* y,
                            const float* z, int64_t stride, int64_t nq, float tol,
                            int32_t* offsets, int32_t* counts, int64_t* total, void* stream) {
  return find_count_any(b, kind, x, y, z, stride, nq, tol, offsets, counts, total, stream, true);
}
int axsq_bvh_find_fill(const axsq_bvh* b, int kind, const double* x, const double* y,
                       const double* z, int64_t stride, int64_t nq, double tol,
                       const int32_t* offsets, const int32_t* counts, int32_t* candidates,
                       void* stream) {
  return find_fill_any(b, kind, x, y, z, stride, nq, tol, offsets, counts, candidates, stream,
                       false);
}
int axsq_bvh_find_fill_f32(const axsq_bvh* b, int kind, const float* x, const float* y,
                           const float* z, int64_t stride, int64_t nq, float tol,
                           const int32_t* offsets, const int32_t* counts, int32_t* candidates,
                           void* stream) {
  return find_fill_any(b, kind, x, y, z, stride, nq, tol, offsets, counts, candidates, stream,
                       true);
}
int axsq_bvh_find_into(const axsq_bvh* b, int kind, const double* x, const double* y,
                       const double* z, int64_t stride, int64_t nq, double tol, int32_t* offsets,
                       int32_t* counts, int32_t* candidates, int64_t capacity, int64_t* total,
                       void* stream) {
  return find_into_any(b, kind, x, y, z, stride, nq, tol, offsets, counts, candidates, capacity,
                       total, stream, false);
}
int axsq_bvh_find_into_f32(const axsq_bvh* b, int kind, const float* x, const float* y,
                           const float* z, int64_t stride, int64_t nq, float tol,
                           int32_t* offsets, int32_t* counts, int32_t* candidates,
                           int64_t capacity, int64_t* total, void* stream) {
  return find_into_any(b, kind, x, y, z, stride, nq, tol, offsets, counts, candidates, capacity,
                       total, stream, true);
}
int axsq_bvh_find(const axsq_bvh* b, int kind, const double* x, const double* y, const double* z,
                  int64_t stride, int64_t nq, double tol, int32_t* offsets, int32_t* counts,
                  int32_t** candidates, int64_t* total, void* stream) {
  return find_alloc_any(b, kind, x, y, z, stride, nq, tol, offsets, counts, candidates, total,
                        stream, false);
}
int axsq_bvh_find_f32(const axsq_bvh* b, int kind, const float* x, const float* y,
                      const float* z, int64_t stride, int64_t nq, float tol, int32_t* offsets,
                      int32_t* counts, int32_t** candidates, int64_t* total, void* stream) {
  return find_alloc_any(b, kind, x, y, z, stride, nq, tol, offsets, counts, candidates, total,
                        stream, true);
}

// ---- SignedDistance -------------------------------------------------------
// First-pass step budget. The long walks (queries nearly equidistant from
// many triangles) grow with the leaf count, and the second pass only pays off
// for walks well above the typical length: 2048 steps at 81,920 leaves,
// scaled by (leaves / 81,920)^(3/4), within [2048, 65536]. It changes speed
// only, never results. AXSQ_SDF_BUDGET overrides it (0: no budget).
static int32_t sdf_budget(int64_t nleaves) {
  int32_t b = (int32_t)std::min(
      65536.0, std::max(2048.0, 2048.0 * std::pow((double)nleaves / 81920.0, 0.75)));
  if (const char* e = getenv("AXSQ_SDF_BUDGET")) b = atoi(e) > 0 ? atoi(e) : 0x7fffffff;
  return b;
}

// AXSQ_SDF_SUSPEND_CAP fixes the number of walk records (speed only: a walk
// over its budget with no record left finishes in the first pass); -1 when
// unset, so the records are sized from the previous call's suspended walks.
static int64_t sdf_suspend_cap() {
  if (const char* e = getenv("AXSQ_SDF_SUSPEND_CAP")) return std::max<int64_t>(atoll(e), 0);
  return -1;
}

// Early walks: the longest walks of a call are remembered (up to this many)
// and walked from the start beside the next call's first pass on the same
// query arrays. They pay where the second pass ends with its longest walk,
// on large trees (a step budget of 8192 or more, about 520k leaves and up:
// the 1.3M-triangle shape-fit loop's forward 273 -> 230 ms); at 81,920
// leaves the second pass is bound by the mass of its walks and the extra
// waves cost the first pass more than they save. AXSQ_SDF_EARLY_WALKS
// overrides the count (0: none). Speed only, never results.
static int64_t sdf_early_walks(int32_t budget) {
  int64_t n = budget >= 8192 ? 64 : 0;
  if (const char* e = getenv("AXSQ_SDF_EARLY_WALKS"))
    n = std::min<int64_t>(std::max<int64_t>(atoll(e), 0), 4096);
  return n;
}

// First pass, calls after the first on the same query arrays: the share (in
// percent) of the previous call's waves with the fewest steps that are
// launched last, so that the first pass ends with short waves rather than
// with whichever waves come last in the query order (its tail on the
// 81,920-triangle headline: active waves fall from 4,000 to 0 over its last
// 2.5 ms). AXSQ_SDF_WAVE_TAIL overrides (0: query order). Speed only.
static int32_t sdf_wave_tail() {
  int32_t v = 20;
  if (const char* e = getenv("AXSQ_SDF_WAVE_TAIL")) v = std::min(std::max(atoi(e), 0), 90);
  return v;
}

// AXSQ_SDF_MIXED_PASS2=fp32: the FP32-traversal form's second pass walks the
// FP32 nodes too (by default it resumes the saved walks on the FP64 nodes).
// Speed only: either way the pruning is exact.
static bool sdf_mixed_pass2_fp32() {
  const char* e = getenv("AXSQ_SDF_MIXED_PASS2");
  return e && strcmp(e, "fp32") == 0;
}

int axsq_sdf_create(axsq_sdf** out, const double* x, const double* y, const double* z,
                    int64_t nverts, const int32_t* tris, int64_t ntris, int watertight,
                    int compute_sign, int fp_mode, void* stream) {
  clear_error();
  if (!out) {
    set_error("axsq_sdf_create: out is NULL");
    return AXSQ_ERR_INVALID_ARG;
  }
  *out = nullptr;
  if (!x || !y || !z || !tris || nverts < 3 || ntris < 1 || nverts > 0x7fffffffll ||
      ntris > 0x3fffffffll) {
    set_error("axsq_sdf_create: need a triangle mesh with >= 3 vertices and >= 1 triangle");
    return AXSQ_ERR_INVALID_ARG;
  }
  if (fp_mode != AXSQ_FP64 && fp_mode != AXSQ_FP32_TRAVERSAL) {
    set_error("axsq_sdf_create: unknown fp_mode %d", fp_mode);
    return AXSQ_ERR_INVALID_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  axsq_sdf* h = new axsq_sdf();
  hipGetDevice(&h->device);
  h->nverts = nverts;
  h->ntris = ntris;
  h->nconn = 3 * ntris;
  h->watertight = watertight ? 1 : 0;
  h->compute_sign = compute_sign ? 1 : 0;
  h->fp_mode = fp_mode;
  h->bvh = new axsq_bvh();
  h->bvh->device = h->device;
  const char* env = getenv("AXSQ_SORT_QUERIES");
  h->sort_queries = !(env && env[0] == '0');
  h->budget = sdf_budget(ntris);
  h->susp_cap_fixed = sdf_suspend_cap();
  h->mixed_pass2_fp32 = sdf_mixed_pass2_fp32();
  h->early_walks = sdf_early_walks(h->budget);
  h->wave_tail = sdf_wave_tail();
  int rc = AXSQ_OK;
  auto fail = [&](int code) {
    axsq_sdf_destroy(h);
    return code;
  };
  if (hipMalloc((void**)&h->vx, sizeof(double) * nverts) != hipSuccess ||
      hipMalloc((void**)&h->vy, sizeof(double) * nverts) != hipSuccess ||
      hipMalloc((void**)&h->vz, sizeof(double) * nverts) != hipSuccess ||
      hipMalloc((void**)&h->conn, sizeof(int32_t) * 3 * ntris) != hipSuccess ||
      hipMalloc((void**)&h->stats_dev, sizeof(int64_t) * 2) != hipSuccess) {
    set_error("axsq_sdf_create: out of device memory");
    return fail(AXSQ_ERR_NOMEM);
  }
  if (hipMemcpyAsync(h->vx, x, sizeof(double) * nverts, hipMemcpyDefault, s) != hipSuccess ||
      hipMemcpyAsync(h->vy, y, sizeof(double) * nverts, hipMemcpyDefault, s) != hipSuccess ||
      hipMemcpyAsync(h->vz, z, sizeof(double) * nverts, hipMemcpyDefault, s) != hipSuccess ||
      hipMemcpyAsync(h->conn, tris, sizeof(int32_t) * 3 * ntris, hipMemcpyDefault, s) !=
          hipSuccess) {
    set_error("axsq_sdf_create: mesh copy failed");
    return fail(AXSQ_ERR_HIP);
  }
  rc = sdf_check_conn(h, s);
  if (rc != AXSQ_OK) return fail(rc);
  rc = sdf_setup(h, s);
  if (rc != AXSQ_OK) return fail(rc);
  rc = sdf_build_v2t(h, s);
  if (rc != AXSQ_OK) return fail(rc);
  *out = h;
  return AXSQ_OK;
}

int axsq_sdf_create_cells(axsq_sdf** out, const double* x, const double* y, const double* z,
                          int64_t nverts, const int32_t* conn, const int32_t* cell_offsets,
                          int64_t ncells, int watertight, int compute_sign, int fp_mode,
                          void* stream) {
  clear_error();
  if (!out) {
    set_error("axsq_sdf_create_cells: out is NULL");
    return AXSQ_ERR_INVALID_ARG;
  }
  *out = nullptr;
  if (!x || !y || !z || !conn || !cell_offsets || nverts < 3 || ncells < 1 ||
      nverts > 0x7fffffffll || ncells > 0x3fffffffll) {
    set_error("axsq_sdf_create_cells: need a mesh with >= 3 vertices and >= 1 cell");
    return AXSQ_ERR_INVALID_ARG;
  }
  if (fp_mode != AXSQ_FP64 && fp_mode != AXSQ_FP32_TRAVERSAL) {
    set_error("axsq_sdf_create_cells: unknown fp_mode %d", fp_mode);
    return AXSQ_ERR_INVALID_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  std::vector<int32_t> off((size_t)ncells + 1);
  if (hipMemcpyAsync(off.data(), cell_offsets, sizeof(int32_t) * off.size(), hipMemcpyDefault,
                     s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error("axsq_sdf_create_cells: cannot read the cell offsets");
    return AXSQ_ERR_HIP;
  }
  bool all_tris = off[0] == 0;
  for (int64_t c = 0; c < ncells; ++c) {
    const int32_t n = off[c + 1] - off[c];
    if (off[c] < 0 || (n != 3 && n != 4)) {
      set_error("axsq_sdf_create_cells: cell %lld has %d nodes (triangles and quads only)",
                (long long)c, n);
      return AXSQ_ERR_INVALID_ARG;
    }
    all_tris = all_tris && (off[c] == 3 * c) && n == 3;
  }
  if (all_tris)  // a triangle mesh: the two-pass path
    return axsq_sdf_create(out, x, y, z, nverts, conn + 0, ncells, watertight, compute_sign,
                           fp_mode, stream);
  axsq_sdf* h = new axsq_sdf();
  hipGetDevice(&h->device);
  h->nverts = nverts;
  h->ntris = ncells;
  h->cells = true;
  h->nconn = off[(size_t)ncells];
  h->watertight = watertight ? 1 : 0;
  h->compute_sign = compute_sign ? 1 : 0;
  h->fp_mode = AXSQ_FP64;  // the cell walk runs on FP64 nodes
  h->budget = sdf_budget(ncells);
  h->susp_cap_fixed = sdf_suspend_cap();
  h->bvh = new axsq_bvh();
  h->bvh->device = h->device;
  const char* env = getenv("AXSQ_SORT_QUERIES");
  h->sort_queries = !(env && env[0] == '0');
  auto fail = [&](int code) {
    axsq_sdf_destroy(h);
    return code;
  };
  if (hipMalloc((void**)&h->vx, sizeof(double) * nverts) != hipSuccess ||
      hipMalloc((void**)&h->vy, sizeof(double) * nverts) != hipSuccess ||
      hipMalloc((void**)&h->vz, sizeof(double) * nverts) != hipSuccess ||
      hipMalloc((void**)&h->conn, sizeof(int32_t) * h->nconn) != hipSuccess ||
      hipMalloc((void**)&h->cell_off, sizeof(int32_t) * off.size()) != hipSuccess ||
      hipMalloc((void**)&h->stats_dev, sizeof(int64_t) * 2) != hipSuccess) {
    set_error("axsq_sdf_create_cells: out of device memory");
    return fail(AXSQ_ERR_NOMEM);
  }
  if (hipMemcpyAsync(h->vx, x, sizeof(double) * nverts, hipMemcpyDefault, s) != hipSuccess ||
      hipMemcpyAsync(h->vy, y, sizeof(double) * nverts, hipMemcpyDefault, s) != hipSuccess ||
      hipMemcpyAsync(h->vz, z, sizeof(double) * nverts, hipMemcpyDefault, s) != hipSuccess ||
      hipMemcpyAsync(h->conn, conn, sizeof(int32_t) * h->nconn, hipMemcpyDefault, s) !=
          hipSuccess ||
      hipMemcpyAsync(h->cell_off, off.data(), sizeof(int32_t) * off.size(),
                     hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error("axsq_sdf_create_cells: mesh copy failed");
    return fail(AXSQ_ERR_HIP);
  }
  int rc = sdf_check_conn(h, s);
  if (rc != AXSQ_OK) return fail(rc);
  rc = sdf_setup(h, s);
  if (rc != AXSQ_OK) return fail(rc);
  *out = h;
  return AXSQ_OK;
}

int axsq_sdf_destroy(axsq_sdf* h) {
  if (!h) return AXSQ_OK;
  hipFree(h->vx);
  hipFree(h->vy);
  hipFree(h->vz);
  hipFree(h->conn);
  hipFree(h->cell_off);
  hipFree(h->tris);
  hipFree(h->early_q);
  hipFree(h->skip);
  hipFree(h->wsteps);
  if (h->early_done) hipEventDestroy(h->early_done);
  hipFree(h->v2t_start);
  hipFree(h->v2t_end);
  hipFree(h->v2t_slot);
  hipFree(h->stats_dev);
  if (h->bvh) {
    bvh_release(h->bvh);
    delete h->bvh;
  }
  delete h;
  return AXSQ_OK;
}

int axsq_sdf_mesh_bounds(const axsq_sdf* h, double lo[3], double hi[3]) {
  clear_error();
  if (!h || !lo || !hi) {
    set_error("axsq_sdf_mesh_bounds: NULL argument");
    return AXSQ_ERR_INVALID_ARG;
  }
  for (int d = 0; d < 3; ++d) {
    lo[d] = h->domain[d];
    hi[d] = h->domain[3 + d];
  }
  return AXSQ_OK;
}

int axsq_sdf_update_vertices(axsq_sdf* h, const double* x, const double* y, const double* z,
                             void* stream) {
  clear_error();
  if (!h || !x || !y || !z) {
    set_error("axsq_sdf_update_vertices: NULL argument");
    return AXSQ_ERR_INVALID_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const size_t nb = sizeof(double) * h->nverts;
  AXSQ_CHECK_HIP(hipMemcpyAsync(h->vx, x, nb, hipMemcpyDefault, s));
  AXSQ_CHECK_HIP(hipMemcpyAsync(h->vy, y, nb, hipMemcpyDefault, s));
  AXSQ_CHECK_HIP(hipMemcpyAsync(h->vz, z, nb, hipMemcpyDefault, s));
  return sdf_setup(h, s);
}

int axsq_sdf_eval(const axsq_sdf* h, const double* x, const double* y, const double* z,
                  int64_t stride, int64_t nq, int fp_mode, double* dist, double* cp,
                  double* normal, int32_t* elem, void* stream) {
  clear_error();
  if (!h) {
    set_error("axsq_sdf_eval: handle is NULL");
    return AXSQ_ERR_STATE;
  }
  if (nq < 0 || nq > 0x7fffffffll || stride < 1) {
    set_error("axsq_sdf_eval: bad query count %lld or stride %lld", (long long)nq,
              (long long)stride);
    return AXSQ_ERR_INVALID_ARG;
  }
  if (nq > 0 && (!x || !y || !z || !dist)) {
    set_error("axsq_sdf_eval: NULL query or distance pointer");
    return AXSQ_ERR_INVALID_ARG;
  }
  if (fp_mode != AXSQ_FP64 && fp_mode != AXSQ_FP32_TRAVERSAL) {
    set_error("axsq_sdf_eval: unknown fp_mode %d", fp_mode);
    return AXSQ_ERR_INVALID_ARG;
  }
  if (fp_mode == AXSQ_FP32_TRAVERSAL &&
      (!h->bvh->nodes32 || h->bvh->fp_mode != AXSQ_FP32_TRAVERSAL)) {
    set_error("axsq_sdf_eval: FP32 traversal needs a handle created with AXSQ_FP32_TRAVERSAL");
    return AXSQ_ERR_STATE;
  }
  if (nq == 0) return AXSQ_OK;
  hipStream_t s = (hipStream_t)stream;
  const bool stats = h->stats && fp_mode == AXSQ_FP64;
  if (stats) AXSQ_TRY(fill_i32(reinterpret_cast<int32_t*>(h->stats_dev), 0, 4, s));
  AXSQ_TRY(sdf_eval_launch(h, x, y, z, stride, nq, fp_mode, dist, cp, normal, elem, stats, s));
  AXSQ_TRY(finish(h->bvh->status, s));
  if (stats) {
    int64_t st[2];
    AXSQ_CHECK_HIP(hipMemcpy(st, h->stats_dev, sizeof(st), hipMemcpyDeviceToHost));
    const_cast<axsq_sdf*>(h)->last_visits = st[0];
    const_cast<axsq_sdf*>(h)->last_leaves = st[1];
  }
  return AXSQ_OK;
}

int axsq_sdf_eval_point(const axsq_sdf* h, const double q[3], int fp_mode, double* dist,
                        double cp[3], double normal[3]) {
  clear_error();
  if (!h || !q || !dist) {
    set_error("axsq_sdf_eval_point: NULL argument");
    return AXSQ_ERR_INVALID_ARG;
  }
  Workspace& ws = workspace();
  AXSQ_TRY(ws.one.ensure(sizeof(double) * 10));
  double* d = ws.one.as<double>();  // query, distance, closest point, normal
  hipStream_t s = nullptr;
  AXSQ_CHECK_HIP(hipMemcpyAsync(d, q, sizeof(double) * 3, hipMemcpyHostToDevice, s));
  AXSQ_TRY(axsq_sdf_eval(h, d, d + 1, d + 2, 3, 1, fp_mode, d + 3, d + 4, d + 7, nullptr, s));
  double out[7];
  AXSQ_CHECK_HIP(hipMemcpy(out, d + 3, sizeof(out), hipMemcpyDeviceToHost));
  *dist = out[0];
  for (int k = 0; k < 3; ++k) {
    if (cp) cp[k] = out[1 + k];
    if (normal) normal[k] = out[4 + k];
  }
  return AXSQ_OK;
}

int axsq_sdf_backward(const axsq_sdf* h, const double* x, const double* y, const double* z,
                      int64_t stride, int64_t nq, const double* dist, const int32_t* elem,
                      const float* gout, double* grad_q, double* grad_v, void* stream) {
  clear_error();
  if (!h) {
    set_error("axsq_sdf_backward: handle is NULL");
    return AXSQ_ERR_STATE;
  }
  if (nq < 0 || nq > 0x7fffffffll || stride < 1) {
    set_error("axsq_sdf_backward: bad query count or stride");
    return AXSQ_ERR_INVALID_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  if (nq == 0) {
    if (grad_v) AXSQ_TRY(fill_i32(reinterpret_cast<int32_t*>(grad_v), 0, 6 * h->nverts, s));
    return finish(nullptr, s);
  }
  if (!x || !y || !z || !dist || !elem || !gout) {
    set_error("axsq_sdf_backward: NULL input");
    return AXSQ_ERR_INVALID_ARG;
  }
  if (h->cells) {
    set_error("axsq_sdf_backward: the adjoint is for triangle meshes (this one has quad cells)");
    return AXSQ_ERR_INVALID_ARG;
  }
  AXSQ_TRY(sdf_backward_launch(h, x, y, z, stride, nq, dist, elem, gout, grad_q, grad_v, s));
  return finish(h->bvh->status, s);
}

// ---- measurement ---------------------------------------------------------
int axsq_prof_enable(int on) {
  g_prof_on = on != 0;
  return AXSQ_OK;
}

int axsq_prof_reset(void) {
  prof_flush();
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_totals.clear();
  return AXSQ_OK;
}

int axsq_prof_get(const char* name, int64_t* launches, double* total_ms) {
  clear_error();
  if (!name || !launches || !total_ms) {
    set_error("axsq_prof_get: NULL argument");
    return AXSQ_ERR_INVALID_ARG;
  }
  prof_flush();
  std::lock_guard<std::mutex> g(g_prof_mu);
  auto it = g_totals.find(name);
  *launches = (it == g_totals.end()) ? 0 : it->second.first;
  *total_ms = (it == g_totals.end()) ? 0.0 : it->second.second;
  return AXSQ_OK;
}

int axsq_bvh_stats_enable(axsq_bvh* b, int on) {
  if (!b) return AXSQ_ERR_INVALID_ARG;
  b->stats = on != 0;
  return AXSQ_OK;
}

int axsq_bvh_stats_get(const axsq_bvh* b, int64_t* visits) {
  if (!b || !visits) return AXSQ_ERR_INVALID_ARG;
  *visits = b->last_visits;
  return AXSQ_OK;
}

int axsq_sdf_stats_enable(axsq_sdf* h, int on) {
  if (!h) return AXSQ_ERR_INVALID_ARG;
  h->stats = on != 0;
  return AXSQ_OK;
}

int axsq_sdf_stats_get(const axsq_sdf* h, int64_t* visits, int64_t* leaves) {
  if (!h || !visits || !leaves) return AXSQ_ERR_INVALID_ARG;
  *visits = h->last_visits;
  *leaves = h->last_leaves;
  return AXSQ_OK;
}

int axsq_sdf_last_suspended(const axsq_sdf* h, int64_t* count) {
  if (!h || !count) return AXSQ_ERR_INVALID_ARG;
  *count = h->last_suspended;
  return AXSQ_OK;
}

int axsq_sdf_last_early(const axsq_sdf* h, int64_t* count) {
  if (!h || !count) return AXSQ_ERR_INVALID_ARG;
  *count = h->last_early;
  return AXSQ_OK;
}

}  // extern "C"
