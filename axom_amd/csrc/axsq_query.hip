// axsq_query.hip -- BVH candidate search (findPoints / findBoundingBoxes / findRays).
//
// Reference: LinearBVH::findCandidatesImpl (src/axom/spin/policy/LinearBVH.hpp:273-402),
// GPU form: count pass, exclusive scan, fill pass (:296-367). One lane per
// query; candidates of query i are written in traversal order at offsets[i],
// at most counts[i] of them (the fill raises kStatusFillMismatch in the status word
// when its walk finds another number, so a caller whose inputs changed after
// the count gets an error and no write past its buffer). In the one-call
// search the count pass also keeps each query's first kstash candidate ids and
// lists the queries that have more, so the fill copies the kept ids and
// searches again only for the listed queries. The walk runs on the vote
// schedule (a lane's order of node visits and leaf actions is the one-lane
// walk's, so the candidate order is too): 11% faster for boxes and 34% for
// rays than the loop-by-loop schedule on 100,000 boxes, the same for points.
//
// FP32 traversal: nodes hold FP32 boxes rounded outward, so every predicate
// evaluated on them passes whenever it passes on the FP64 box (points and
// boxes as FP32 intervals of the exact query; rays with the FP64 slab test,
// which is monotone under box growth). Each reached leaf is confirmed on its
// FP64 box, so the candidate lists are identical to the FP64 traversal's,
// order included.
#include "axsq_internal.hpp"
#include "axsq_traverse.hpp"

namespace axsq {

namespace {

struct FindArgs {
  const Node64* nodes;
  const Node32* nodes32;
  const double* leaf_boxes;
  const int32_t* leaf_ids;
  const void* x;  // double queries, or float ones for a float tree (f32)
  const void* y;
  const void* z;
  int64_t stride;
  int f32;  // a float tree: float queries, the ray test in float
  int64_t nq;
  int dims;  // 2: queries are 2-D (z = 0)
  int32_t ninner;
  double tol;
  int32_t* counts;           // count pass: output
  const int32_t* fill_cnt;   // fill: each query's room, the counts its count pass wrote
  const int32_t* offsets;
  int32_t* cands;
  int32_t* status;
  int32_t* stash;      // count pass: kstash ids per query (or NULL)
  int kstash;
  int32_t* stash_cnt;  // count pass: a copy of the counts
  int32_t* over;       // count pass: queries with more than kstash; fill: those to search
  uint32_t* nover;
  unsigned long long* visits;  // count pass: sum of the walks' inner-node visits (or NULL)
};

enum { kCountPass = 0, kFillAll = 1, kFillListed = 2 };

// Predicates of BVH.hpp:499-501 (contains), :558-560 (intersectsWith),
// :529-532 (intersect_ray with the BVH tolerance).
template <int KIND>
__device__ __forceinline__ bool pred64(const double* prim, double tol, const double b[6],
                                       int f32 = 0, int dims = 3) {
  if (KIND == AXSQ_QUERY_POINTS) {
    return !(prim[0] < b[0] || prim[0] > b[3] || prim[1] < b[1] || prim[1] > b[4] ||
             prim[2] < b[2] || prim[2] > b[5]);
  } else if (KIND == AXSQ_QUERY_BOXES) {
    return (prim[3] >= b[0] && prim[0] <= b[3]) && (prim[4] >= b[1] && prim[1] <= b[4]) &&
           (prim[5] >= b[2] && prim[2] <= b[5]);
  } else {
    return f32 ? ray_box_f32(prim, b, (float)tol, dims) : ray_box(prim, b, tol);
  }
}

// Superset predicates on outward-rounded FP32 boxes. lo/hi bracket the query
// (points: the point; boxes: min rounded down / max rounded up).
template <int KIND>
__device__ __forceinline__ bool pred32(const double* prim, const float* lo, const float* hi,
                                       double tol, const float b[6], int f32 = 0, int dims = 3) {
  if (KIND == AXSQ_QUERY_RAYS) {
    const double bd[6] = {b[0], b[1], b[2], b[3], b[4], b[5]};
    return f32 ? ray_box_f32(prim, bd, (float)tol, dims) : ray_box(prim, bd, tol);
  } else if (KIND == AXSQ_QUERY_POINTS) {
    // written as pred64's contains(): a NaN coordinate passes, as there
    return !(hi[0] < b[0] || lo[0] > b[3] || hi[1] < b[1] || lo[1] > b[4] || hi[2] < b[2] ||
             lo[2] > b[5]);
  } else {
    return (hi[0] >= b[0] && lo[0] <= b[3]) && (hi[1] >= b[1] && lo[1] <= b[4]) &&
           (hi[2] >= b[2] && lo[2] <= b[5]);
  }
}

// Stack entries per lane in LDS: a point's or box's walk pushes one entry per
// node where both children hold the query, rarely many, so 16 (deeper ones in
// the private array) let 8 waves per SIMD share a CU's LDS instead of 5
// (findPoints 1M x 100k 0.486 -> 0.461 ms; the hex candidate search 0.86 ->
// 0.77 ms). Ray walks keep 32.
template <int KIND>
constexpr int find_lds_stack() { return KIND == AXSQ_QUERY_RAYS ? kLdsStack : 16; }

template <int KIND, int MODE, bool MIXED>
__global__ __launch_bounds__(kBlock) void k_find(FindArgs a) {
  constexpr bool FILL = MODE != kCountPass;
  constexpr int kLds = find_lds_stack<KIND>();
  __shared__ int32_t s_stack[stack_lds_words<kLds>(kBlock)];
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = j < a.nq;
  if (FILL && !live) return;  // the count pass keeps whole waves for its ballot
  int32_t count = 0;
  int64_t i = 0;
  if (live) {
    i = (MODE == kFillListed) ? (int64_t)a.over[j] : j;
    lds_i32* stk = lane_stack_base<kLds>(s_stack);
    double prim[6];
    // element k of query array p (float queries widen exactly)
    auto ld = [&](const void* p, int64_t k) -> double {
      return a.f32 ? (double)static_cast<const float*>(p)[k] : static_cast<const double*>(p)[k];
    };
    if (KIND == AXSQ_QUERY_POINTS) {
      prim[0] = ld(a.x, i * a.stride);
      prim[1] = ld(a.y, i * a.stride);
      prim[2] = (a.dims == 2) ? 0.0 : ld(a.z, i * a.stride);
      prim[3] = prim[4] = prim[5] = 0.0;
    } else if (a.dims == 2) {  // box: min x y, max x y; ray: origin x y, direction x y
      const int64_t o = i * a.stride;
      prim[0] = ld(a.x, o);
      prim[1] = ld(a.x, o + 1);
      prim[2] = 0.0;
      prim[3] = ld(a.x, o + 2);
      prim[4] = ld(a.x, o + 3);
      prim[5] = 0.0;
    } else {
      for (int k = 0; k < 6; ++k) prim[k] = ld(a.x, i * a.stride + k);
    }
    const int32_t offset = FILL ? a.offsets[i] : 0;
    const int32_t room = FILL ? a.fill_cnt[i] : 0;
    int32_t* const stash = (!FILL && a.stash) ? a.stash + i * a.kstash : nullptr;
    auto emit = [&](int32_t pos) {
      if (FILL) {
        if (count < room) a.cands[offset + count] = a.leaf_ids[pos];
      } else {
        if (stash && count < a.kstash) stash[count] = a.leaf_ids[pos];
      }
      ++count;
    };
    auto no_pref = [](const auto*, const auto*) { return false; };
    int32_t v;
    if (!MIXED) {
      // the walk tests a child box only when its validity bit is set, so an
      // invalid box (min > max on some axis) is never entered
      v = traverse_voted<Node64, double, kLds>(
          a.nodes, a.ninner, stk,
          [&](const double* b) { return pred64<KIND>(prim, a.tol, b, a.f32, a.dims); }, no_pref,
          emit);
    } else {
      float lo[3], hi[3];
      const int hoff = (KIND == AXSQ_QUERY_BOXES) ? 3 : 0;
      for (int d = 0; d < 3; ++d) {
        lo[d] = __double2float_rd(prim[d]);
        hi[d] = __double2float_ru(prim[hoff + d]);
      }
      v = traverse_voted<Node32, float, kLds>(
          a.nodes32, a.ninner, stk,
          [&](const float* b) { return pred32<KIND>(prim, lo, hi, a.tol, b, a.f32, a.dims); },
          no_pref,
          [&](int32_t pos) {
            double lb[6];
            const double2* p = reinterpret_cast<const double2*>(a.leaf_boxes + 6 * (int64_t)pos);
            const double2 c0 = p[0], c1 = p[1], c2 = p[2];
            lb[0] = c0.x; lb[1] = c0.y; lb[2] = c1.x; lb[3] = c1.y; lb[4] = c2.x; lb[5] = c2.y;
            if (valid6(lb) && pred64<KIND>(prim, a.tol, lb, a.f32, a.dims)) emit(pos);
          });
    }
    if (v < 0) atomicOr(a.status, kStatusStack);
    if (!FILL && a.visits && v > 0) atomicAdd(a.visits, (unsigned long long)v);
    if (FILL && count != room) atomicOr(a.status, kStatusFillMismatch);
    if (!FILL) {
      a.counts[i] = count;
      if (stash) a.stash_cnt[i] = count;
    }
  }
  if (!FILL && a.stash) {  // queries with more than kstash: one list slot per wave
    const bool more = live && count > a.kstash;
    const uint64_t m = __ballot(more);
    if (m) {
      const int lane = (int)__lane_id();
      const int leader = __ffsll((unsigned long long)m) - 1;
      uint32_t base = 0;
      if (lane == leader) base = atomicAdd(a.nover, (uint32_t)__popcll(m));
      base = __shfl(base, leader);
      const uint64_t below = (lane == 0) ? 0ull : (m & ((~0ull) >> (64 - lane)));
      if (more) a.over[base + __popcll(below)] = (int32_t)i;
    }
  }
}

// The fill pass for queries with at most kstash (<= 64) candidates: the ids
// the count pass kept, at offsets[i]. A wave whose queries have at most 4 ids
// each copies one query per lane; otherwise it copies its 64 queries' ids one
// query at a time, lane l taking id l, so each load and store is one
// contiguous run, with eight queries' loads issued before their stores.
__global__ __launch_bounds__(kBlock) void k_stash_copy(const int32_t* __restrict__ stash,
                                                        int kstash,
                                                        const int32_t* __restrict__ stash_cnt,
                                                        const int32_t* __restrict__ offsets,
                                                        int64_t nq, int32_t* __restrict__ cands) {
  const int lane = (int)__lane_id();
  const int64_t base = (int64_t)blockIdx.x * blockDim.x + threadIdx.x - lane;  // the wave's first
  if (base >= nq) return;  // the whole wave
  const int64_t mine = base + lane;
  const int32_t c_me = (mine < nq) ? stash_cnt[mine] : 0;
  const int32_t o_me = (mine < nq) ? offsets[mine] : 0;
  if (__ballot(c_me > 4 && c_me <= kstash) == 0) {  // few ids everywhere: a lane per query
    if (c_me <= kstash)
      for (int32_t k = 0; k < c_me; ++k) cands[o_me + k] = stash[mine * kstash + k];
    return;
  }
  for (int j0 = 0; j0 < 64; j0 += 8) {
    int32_t v[8], cc[8], oo[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      cc[u] = __shfl(c_me, j0 + u);
      oo[u] = __shfl(o_me, j0 + u);
      const bool take = cc[u] <= kstash && lane < cc[u];
      v[u] = take ? stash[(base + j0 + u) * kstash + lane] : 0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (cc[u] <= kstash && lane < cc[u]) cands[oo[u] + lane] = v[u];
  }
}

template <int MODE, bool MIXED>
hipError_t launch_kind(int kind, const FindArgs& a, hipStream_t s) {
  const unsigned g = grid_for(a.nq);
  switch (kind) {
    case AXSQ_QUERY_POINTS:
      k_find<AXSQ_QUERY_POINTS, MODE, MIXED><<<g, kBlock, 0, s>>>(a);
      break;
    case AXSQ_QUERY_BOXES:
      k_find<AXSQ_QUERY_BOXES, MODE, MIXED><<<g, kBlock, 0, s>>>(a);
      break;
    default:
      k_find<AXSQ_QUERY_RAYS, MODE, MIXED><<<g, kBlock, 0, s>>>(a);
      break;
  }
  return hipGetLastError();
}

template <bool MIXED>
hipError_t launch_mode(int mode, int kind, const FindArgs& a, hipStream_t s) {
  if (mode == kCountPass) return launch_kind<kCountPass, MIXED>(kind, a, s);
  if (mode == kFillAll) return launch_kind<kFillAll, MIXED>(kind, a, s);
  return launch_kind<kFillListed, MIXED>(kind, a, s);
}

}  // namespace

// mode 0: count (and keep the first kstash ids when stash is given); 1: fill
// every query; 2: fill the nq queries listed in `over`.
int find_launch(const axsq_bvh* b, int kind, int mode, const void* x, const void* y,
                const void* z, int64_t stride, int64_t nq, double tol, int32_t* counts,
                const int32_t* fill_cnt, const int32_t* offsets, int32_t* cands, int32_t* stash,
                int kstash, int32_t* stash_cnt, int32_t* over, uint32_t* nover, hipStream_t s) {
  FindArgs a;
  a.nodes = b->nodes;
  a.nodes32 = b->nodes32;
  a.leaf_boxes = b->leaf_boxes;
  a.leaf_ids = b->leaf_ids;
  a.x = x;
  a.y = y;
  a.z = z;
  a.stride = stride;
  a.f32 = b->f32 ? 1 : 0;
  a.nq = nq;
  a.dims = b->dims;
  a.ninner = (int32_t)(b->n - 1);
  a.tol = tol;
  a.counts = counts;
  a.fill_cnt = fill_cnt;
  a.offsets = offsets;
  a.cands = cands;
  a.status = b->status;
  a.stash = stash;
  a.kstash = kstash;
  a.stash_cnt = stash_cnt;
  a.over = over;
  a.nover = nover;
  a.visits = (mode == kCountPass && b->stats) ? b->find_visits : nullptr;
  const bool mixed = (b->fp_mode == AXSQ_FP32_TRAVERSAL) && b->nodes32;
  ProfRange pr(mode == kCountPass ? "bvh_find_count" : "bvh_find_fill", s);
  const hipError_t e = mixed ? launch_mode<true>(mode, kind, a, s) : launch_mode<false>(mode, kind, a, s);
  pr.stop();
  if (e != hipSuccess) {
    set_error("find kernel launch failed: %s", hipGetErrorString(e));
    return AXSQ_ERR_HIP;
  }
  return AXSQ_OK;
}

int stash_copy_launch(const int32_t* stash, int kstash, const int32_t* stash_cnt,
                      const int32_t* offsets, int64_t nq, int32_t* cands, hipStream_t s) {
  if (nq <= 0) return AXSQ_OK;
  ProfRange pr("bvh_find_copy", s);
  k_stash_copy<<<grid_for(nq), kBlock, 0, s>>>(stash, kstash, stash_cnt, offsets, nq, cands);
  AXSQ_CHECK_HIP(hipGetLastError());
  return AXSQ_OK;
}

}  // namespace axsq
