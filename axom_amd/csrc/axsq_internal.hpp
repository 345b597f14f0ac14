// axsq_internal.hpp -- host-side plumbing shared by the libaxsq translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "axsq.h"
#include "axsq_device.hpp"

namespace axsq {

// ---- errors ---------------------------------------------------------------
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();

#define AXSQ_CHECK_HIP(call)                                                   \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::axsq::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                   \
      return AXSQ_ERR_HIP;                                                     \
    }                                                                          \
  } while (0)

#define AXSQ_TRY(expr)       \
  do {                       \
    int rc_ = (expr);        \
    if (rc_ != AXSQ_OK) return rc_; \
  } while (0)

// Bits of a tree's device status word (axsq_bvh::status), raised by kernels
// and turned into error codes when the call finishes.
enum : int32_t {
  kStatusStack = 1,         // a walk's stack went deeper than kStack entries
  kStatusElem = 2,          // an element id out of range in the adjoint input
  kStatusFillMismatch = 4,  // a fill found another number of candidates than counted
};

// ---- device buffers -------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t b);  // grows (never shrinks); contents are not preserved
  void release();
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

// Scratch owned by the calling host thread (per device), reused across calls.
struct Workspace {
  DevBuf k0, v0, k1, v1, hist, tiles, misc, flag, susp, cnt, redo;
  DevBuf mesh_tri, mesh_box, mesh_int, mesh_slots;  // mesh self-intersection
  DevBuf box2d;                                      // 2-D boxes widened to 3-D
  DevBuf hot;                                        // tables of the longest saved walks
  DevBuf wave_cost, wave_order;                      // first pass: launch order of the waves
  axsq_bvh* mesh_bvh = nullptr;                      // its BVH, rebuilt per call
  // one-call candidate search: the count pass keeps each query's first
  // candidates (find_stash, find_cnt) and lists the queries with more
  // (find_over, their number in find_n), and the fill of the same call
  // copies them instead of searching again
  DevBuf find_stash, find_cnt, find_over, find_n;
  DevBuf one;  // one query point and its outputs (axsq_sdf_eval_point)
  int device = -1;
  void release();
};
Workspace& workspace();  // for the current device

// ---- profiling (HIP events on the launching stream) -----------------------
bool prof_on();
struct ProfRange {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const char* name = nullptr;
  hipStream_t s = nullptr;
  ProfRange(const char* n, hipStream_t st);
  ProfRange(const ProfRange&) = delete;
  ProfRange& operator=(const ProfRange&) = delete;
  ~ProfRange() { stop(); }
  void stop();  // records the end event once; later calls do nothing
};
void prof_flush();  // after a stream sync: folds finished ranges into totals

// Candidate ids a query's count pass keeps for the fill pass: 64 (the copy
// kernel takes at most one wave's width), fewer when that would take more
// than 512 MB (at least 4).
inline int stash_slots(int64_t nq) {
  int k = 64;
  while (k > 4 && (int64_t)k * 4 * nq > (int64_t(512) << 20)) k >>= 1;
  return k;
}

// ---- primitives (axsq_prims.hip) ------------------------------------------
constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;

inline unsigned grid_for(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  return (unsigned)(g < 1 ? 1 : g);
}

int fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s);
// Exclusive scan of n int32 counts into out; *total_dev (int64) receives the sum.
int scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n, int64_t* total_dev,
                       Workspace& ws, hipStream_t s);
// The same scan with the caller's buffer of scan_tiles(n) int64 tile sums.
inline int64_t scan_tiles(int64_t n) {
  const int64_t t = (n + kTile - 1) / kTile;
  return t < 1 ? 1 : t;
}
int scan_exclusive_i32_into(const int32_t* in, int32_t* out, int64_t n, int64_t* sums,
                            int64_t* total_dev, hipStream_t s);
// Stable sort of 2 .. kSmallSortMax (key, val) pairs by the low `bits` (9 to
// 32) bits for the BVH build, in 8-bit passes over tiles of kSmallSortTile
// keys. hists holds one tile-major histogram per pass (256 * hs words,
// hs = small_sort_stride(n) tiles; hist[tile * 256 + digit]): the first
// filled by the caller, the others zero (each pass's scatter adds the next
// pass's counts). The last pass
// writes the sorted keys to keys_last and, at each sorted position p, the
// key's value to vals_last[p] and row vals_last[p] of `rows` (6 doubles) to
// rows_out[p]; keys/vals and their alternates are scratch.
constexpr int64_t kSmallSortTile = 1024;
constexpr int64_t kSmallSortMax = 256 * kSmallSortTile;
inline int32_t small_sort_stride(int64_t n) {
  return (int32_t)((((n + kSmallSortTile - 1) / kSmallSortTile) + 3) & ~int64_t(3));
}
int radix_sort_small_build(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                           uint32_t* vals_alt, int64_t n, int bits, uint32_t* hists,
                           const double* rows, double* rows_out, uint32_t* keys_last,
                           int32_t* vals_last, hipStream_t s);
// Stable ascending sort of (keys, vals) by the low `bits` bits of the key.
// On return the sorted data is in (*keys_out, *vals_out), which point at
// either the input arrays or the alternates.
int radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                     int64_t n, int bits, Workspace& ws, hipStream_t s, uint32_t** keys_out,
                     uint32_t** vals_out);

inline int bits_for(uint64_t max_value) {
  int b = 1;
  while (b < 32 && (max_value >> b) != 0) ++b;
  return b;
}

}  // namespace axsq

// ---- handles -------------------------------------------------------------
struct axsq_bvh {
  int64_t n = 0;        // leaves after padding
  int dims = 3;         // 2: a 2-D tree (boxes held with z = [0, 0]), 16-bit-per-axis codes
  int fp_mode = AXSQ_FP64;
  double scale = 1.000123;
  double bounds[6];     // host copy of the root box
  int device = 0;
  axsq::Node64* nodes = nullptr;    // n-1 FP64 nodes, root = 0
  axsq::Node32* nodes32 = nullptr;  // n-1 FP32 nodes (mixed precision)
  int4* ranges32 = nullptr;         // with nodes32: each node's leaf range and split
                                    // {lo, hi, mid, 0}, so an FP32 walk reads 16 B, not
                                    // the FP64 node's 128-B line
  double* leaf_boxes = nullptr;     // n scaled leaf boxes in leaf order (AoS)
  int32_t* leaf_ids = nullptr;      // n: leaf position -> original id
  uint32_t* mcodes = nullptr;       // n sorted Morton codes
  int32_t* status = nullptr;        // device status word (kStatus* bits)
  int allocator = AXSQ_ALLOC_DEVICE;  // memory of the candidate arrays axsq_bvh_find returns
  bool f32 = false;  // a float tree (FloatType = float): boxes are float values, queries float
  // traversal statistics of the candidate search's count pass
  bool stats = false;
  unsigned long long* find_visits = nullptr;  // device: inner-node visits summed over queries
  int64_t last_visits = 0;                    // host: the last count pass's sum
  // builds of at most kSmallSortMax boxes replay a HIP graph of their
  // launches while every pointer and size they were captured with holds
  hipGraphExec_t build_graph = nullptr;
  std::vector<uint64_t> build_graph_key;
  // pinned host memory the build's bounds kernel writes its parameters to
  // (the bounds and the signed-zero flag), read after the stream sync
  void* host_params = nullptr;       // host pointer
  void* host_params_dev = nullptr;   // the same memory as the device sees it
};

struct axsq_sdf {
  int64_t nverts = 0, ntris = 0;  // ntris: cells (triangles or quads) in cell mode
  bool cells = false;             // some cell is a quad: two triangle records per leaf
  int64_t nconn = 0;              // connectivity length
  int32_t* cell_off = nullptr;    // cell mode: ncells + 1 offsets into conn
  int watertight = 1, compute_sign = 1, fp_mode = AXSQ_FP64;
  int device = 0;
  double* vx = nullptr;
  double* vy = nullptr;
  double* vz = nullptr;
  int32_t* conn = nullptr;
  double domain[6];                 // host copy of the mesh bounds
  axsq_bvh* bvh = nullptr;
  axsq::TriRec* tris = nullptr;     // triangles in BVH leaf order
  int32_t* v2t_start = nullptr;     // per vertex: first incident slot in v2t_slot
  int32_t* v2t_end = nullptr;
  int32_t* v2t_slot = nullptr;      // incident (3*tri + corner), ascending
  bool stats = false;
  int64_t* stats_dev = nullptr;     // [visits, leaves]
  int64_t last_visits = 0, last_leaves = 0;
  bool sort_queries = true;
  int32_t budget = 2048;   // first-pass step budget per query walk
  int64_t last_suspended = 0;  // walks the last eval finished in its second pass
  // records for suspended walks: sized from the walks the previous call
  // suspended (susp_demand, -1 before the first call), or fixed by
  // AXSQ_SDF_SUSPEND_CAP at creation (susp_cap_fixed, -1 when unset)
  int64_t susp_demand = -1;
  int64_t susp_cap_fixed = -1;
  // FP32-traversal form: second pass on the FP32 nodes too (AXSQ_SDF_MIXED_PASS2=fp32
  // at creation); by default it resumes the saved walks on the FP64 nodes
  bool mixed_pass2_fp32 = false;
  // the longest walks of the previous call on the same query arrays (early_key:
  // x, y, z, stride, nq): query indices on the device, walked from the start
  // beside the next such call's first pass; skip marks them for that pass
  int32_t* early_q = nullptr;
  int64_t nearly = 0;
  uint64_t early_key[5] = {0, 0, 0, 0, 0};
  int32_t* skip = nullptr;
  int64_t skip_cap = 0;
  hipEvent_t early_done = nullptr;  // recorded where a call last touches them (and wsteps)
  int64_t last_early = 0;  // walks the last call ran beside its first pass
  int64_t early_walks = 0;  // walks a call remembers for the next (set at creation)
  // first-pass steps of each query of the previous call, at its position in
  // the query order (wave_key: x, y, z, stride, nq of that call); the next call
  // on the same arrays launches the waves of least steps last
  int32_t* wsteps = nullptr;
  int64_t wsteps_cap = 0;
  uint64_t wave_key[5] = {0, 0, 0, 0, 0};
  bool wsteps_valid = false;
  int32_t wave_tail = 0;  // percent of waves launched last (set at creation; 0: off)
};

namespace axsq {
// boxes: AoS of 6 doubles (min xyz, max xyz); dims 2 builds the 2-D tree of
// boxes whose z range is [0, 0]. f32: a float tree over boxes_f32, AoS of
// 2 * dims floats (boxes unused), with fp_mode AXSQ_FP32_TRAVERSAL.
int bvh_build_into(axsq_bvh* b, const double* boxes, int64_t n, double scale, int fp_mode,
                   hipStream_t s, int dims = 3, bool f32 = false,
                   const float* boxes_f32 = nullptr);
void bvh_release(axsq_bvh* b);
// n 2-D boxes (4 doubles) -> 3-D boxes with z range [0, 0] (6 doubles)
int widen_boxes2d(const double* boxes2d, int64_t n, double* out, hipStream_t s);
// Candidate search launches (axsq_query.hip). mode 0: count pass (keeping the
// first kstash ids per query when stash is given, listing the queries with
// more in `over`); 1: fill every query; 2: fill the nq queries listed in
// `over`. A fill writes at most fill_cnt[i] ids for query i and raises
// kStatusFillMismatch when its walk finds another number. x, y, z point at
// doubles, or at floats for a float tree (b->f32).
int find_launch(const axsq_bvh* b, int kind, int mode, const void* x, const void* y,
                const void* z, int64_t stride, int64_t nq, double tol, int32_t* counts,
                const int32_t* fill_cnt, const int32_t* offsets, int32_t* cands, int32_t* stash,
                int kstash, int32_t* stash_cnt, int32_t* over, uint32_t* nover, hipStream_t s);
// The fill for queries with at most kstash candidates, from the kept ids.
int stash_copy_launch(const int32_t* stash, int kstash, const int32_t* stash_cnt,
                      const int32_t* offsets, int64_t nq, int32_t* cands, hipStream_t s);
}  // namespace axsq
