// axsq_sdf.hip -- signed distance to a triangle surface on gfx950.
//
// Forward: SignedDistance::computeDistances (src/axom/quest/SignedDistance.hpp:529-605)
// -- one lane per query runs the nearer-first traversal with the stateful
// pruning predicate squared_distance(p, box) <= minSqDist, checks each leaf
// triangle (checkCandidate :637-737), then decides the sign (:580-592,
// :741-763). Queries are processed in Morton order of their position (a
// permutation applied on input and undone on output) so neighbouring lanes
// walk neighbouring parts of the tree.
//
// FP32 traversal: node boxes are FP32, rounded outward. The box test is
// sum_d gap_d^2 <= t, with gap_d taken against the FP32 interval
// [rd(q_d), ru(q_d)] of the exact query and
// t = ru(minSqDist * (1 + 1e-6) + 1e-36). Each FP32 operation in
// round-to-nearest is off by at most 2^-24 relative (plus a subnormal
// absolute error far below 1e-36); five of them stay far below 1e-6, and
// the FP32 box contains the FP64 box, so every node the FP64 predicate keeps
// is also kept: the nearest triangle is always checked. Triangle tests, the
// closest point, the pseudo-normal and the sign are FP64.
//
// Adjoint (no reference counterpart): see axsq_sdf_backward in axsq.h.
#include <algorithm>
#include <cstring>
#include <type_traits>

#include "axsq_internal.hpp"
#include "axsq_traverse.hpp"

#ifndef AXSQ_SDF_BLOCK
#define AXSQ_SDF_BLOCK 64  // one wave per block: a finished wave frees its LDS at once
#endif
#ifndef AXSQ_SDF_P1_ILP
#define AXSQ_SDF_P1_ILP 0  // first pass: branch-free closest point
#endif
#ifndef AXSQ_SDF_VOTE
#define AXSQ_SDF_VOTE 1  // first pass: vote-scheduled walk (0: loop-by-loop)
#endif
// second pass: waves per SIMD, FP64 / FP32-traversal form (the FP32 form,
// with its two box tests, would spill at 3; so would the quad-cell form, with
// two records per memo lane)
#ifndef AXSQ_RESUME_WAVES
#define AXSQ_RESUME_WAVES 3
#endif
#ifndef AXSQ_RESUME_WAVES_MIXED
#define AXSQ_RESUME_WAVES_MIXED 2
#endif
// second pass: the longest saved walks (at most this many, and at most 2 GB
// of tables) get their leaf records and node keys computed beforehand
#ifndef AXSQ_HOT_WALKS
#define AXSQ_HOT_WALKS 32
#endif

#define AXSQ_RESUME_WAVES_ATTR \
  __attribute__((amdgpu_waves_per_eu((MIXED || CELLS) ? AXSQ_RESUME_WAVES_MIXED : AXSQ_RESUME_WAVES)))
#if defined(AXSQ_SDF_WAVES) && AXSQ_SDF_WAVES > 0
#define AXSQ_SDF_WAVES_ATTR __attribute__((amdgpu_waves_per_eu(AXSQ_SDF_WAVES)))
#else
#define AXSQ_SDF_WAVES_ATTR
#endif

namespace axsq {

namespace {

constexpr int kSdfBlock = AXSQ_SDF_BLOCK;

struct SdfArgs {
  const Node64* nodes;
  const Node32* nodes32;
  const int4* ranges32;  // with nodes32: {lo, hi, mid, 0} per node
  const TriRec* tris;
  const double* qx;
  const double* qy;
  const double* qz;
  int64_t qs;
  int64_t nq;
  int32_t ninner;
  const uint32_t* order;
  const int32_t* skip;  // first pass: queries whose walks run elsewhere (nonzero), or NULL
  const int32_t* wave_order;  // first pass: wave of the query order per launched wave, or NULL
  int32_t* wsteps;            // first pass: each walk's steps at its query-order position, or NULL
  int32_t budget;      // first-pass step budget per walk
  int64_t cap;         // capacity of susp
  struct SdfSuspended* susp;
  uint32_t* nsusp;
  int32_t* redo;       // queries to run again from the start (room for nq)
  uint32_t* nredo;
  const uint32_t* susp_order;  // second pass: saved walks in launch order (NULL: as saved)
  int64_t nhot;                // second pass: walks 0 .. nhot-1 of that order read tables
  const void* hot_leaf;        // HotLeaf[nhot][ninner + 1]
  const void* hot_node;        // HotNode<Scalar>[nhot][ninner]
  double dom[6];
  int watertight;
  int compute_sign;
  double* dist;
  double* cp;
  double* normal;
  int32_t* elem;
  int32_t* status;
  int64_t* stats;
};

__device__ __forceinline__ void load_tri(const TriRec* __restrict__ tris, int32_t pos, double T[9]) {
  const double2* p = reinterpret_cast<const double2*>(tris + pos);
  const double2 a = p[0], b = p[1], c = p[2], d = p[3];
  T[0] = a.x; T[1] = a.y; T[2] = b.x; T[3] = b.y; T[4] = c.x; T[5] = c.y;
  T[6] = d.x; T[7] = d.y;
  T[8] = reinterpret_cast<const double*>(p)[8];
}

// MinCandidate (SignedDistance.hpp:159-175) plus the leaf position of the winner.
struct MinCand {
  double sq;
  double p[3];
  double sn[3];
  int type;  // -1 unset, 0 vertex, 1 edge, 2 face (getClosestPointLocType :119-137)
  int32_t pos;
};

// The pseudo-normal term a triangle adds when its closest point is shared:
// code 3 -- an edge: the unit face normal; code 0..2 -- that vertex: the unit
// face normal times the triangle's angle there, or nothing for a degenerate
// triangle (SignedDistance.hpp:700-734, Triangle::degenerate Triangle.hpp:327-330).
// Zero is returned for nothing: adding +0.0 to the running sum changes nothing,
// since that sum starts at +0.0 and a round-to-nearest sum is -0.0 only when
// both terms are.
__device__ __forceinline__ void normal_term(const double T[9], int code, double c[3]) {
  double n[3], un[3];
  tri_normal(T, n);
  if (code == 3) {
    unit3(n, un);
    c[0] = un[0];
    c[1] = un[1];
    c[2] = un[2];
    return;
  }
  c[0] = c[1] = c[2] = 0.0;
  const double area = 0.5 * sqrt(dot3(n[0], n[1], n[2], n[0], n[1], n[2]));
  if (!near_eq(area, 0.0, 1.0e-12)) {
    const double alpha = tri_angle(T, code);
    unit3(n, un);
    c[0] = un[0] * alpha;
    c[1] = un[1] * alpha;
    c[2] = un[2] * alpha;
  }
}

// checkCandidate, SignedDistance.hpp:637-737 (triangle cells), first-pass
// form. Most pseudo-normal terms a walk computes are thrown away by a later
// strictly closer candidate, and in a wave one lane's term costs every lane
// the instructions (a square root, divisions, an arc cosine). So the walk
// only records which triangles contribute since the last reset -- leaf
// position and term code, packed as pos * 4 + code (positions below 2^30) --
// and the terms are computed and added in the recorded order once the walk
// has stopped: the same additions in the same order. The list holds
// kNormList entries; one more sets kListOverflow, and such a walk is run
// again from the start by the second pass, which adds terms as it goes (a
// reset also clears the overflow: the terms before it no longer count).
constexpr int kNormList = 8;
constexpr int32_t kListOverflow = 1 << 16;

__device__ __forceinline__ void check_candidate_deferred(MinCand& m, const double q[3],
                                                         const double T[9], int32_t pos,
                                                         bool compute_normal, lds_i32* list,
                                                         int32_t& cnt) {
  constexpr double EPS = 1e-12;
  double cp[3];
#if AXSQ_SDF_P1_ILP
  const int loc = closest_point_ilp(q, T, EPS, cp);
#else
  double wdummy[3];
  const int loc = closest_point<false>(q, T, EPS, cp, wdummy);
#endif
  const double sq = dot3(cp[0] - q[0], cp[1] - q[1], cp[2] - q[2], cp[0] - q[0], cp[1] - q[1],
                         cp[2] - q[2]);
  const int type = (loc < 0) ? 1 : (loc < 3) ? 0 : 2;
  const bool shared = (type != 2);
  const double e0 = m.p[0] - cp[0], e1 = m.p[1] - cp[1], e2 = m.p[2] - cp[2];
  const bool same = near_eq(dot3(e0, e1, e2, e0, e1, e2), 0., EPS);
  bool upd;
  if (sq < m.sq) {
    const bool clear = !shared || (m.type != type) || !same;
    m.sq = sq;
    m.p[0] = cp[0];
    m.p[1] = cp[1];
    m.p[2] = cp[2];
    m.type = type;
    m.pos = pos;
    if (compute_normal && clear) cnt = 0;
    upd = compute_normal && shared;
  } else {
    upd = compute_normal && shared && (m.type == type) && same;
  }
  if (upd) {
    if (cnt < kNormList) {
      list[cnt * kWaveLanes] = (int32_t)((uint32_t)pos * 4u + (uint32_t)(type == 1 ? 3 : loc));
      ++cnt;
    } else {
      cnt |= kListOverflow;
    }
  }
}

// Adds the recorded terms, in order, to sn. `stride` is the distance between
// consecutive entries (kWaveLanes for a lane's LDS column, 1 for a record).
template <class P>
__device__ __forceinline__ void add_terms(double sn[3], const TriRec* __restrict__ tris, P list,
                                          int32_t stride, int32_t cnt) {
  for (int32_t k = 0; k < cnt; ++k) {
    const uint32_t e = (uint32_t)list[k * stride];
    double T[9], c[3];
    load_tri(tris, (int32_t)(e >> 2), T);
    normal_term(T, (int)(e & 3u), c);
    sn[0] += c[0];
    sn[1] += c[1];
    sn[2] += c[2];
  }
}

// Surface normal (getSurfaceNormal, SignedDistance.hpp:741-749), sign
// (:580-592) and the outputs of query i.
__device__ __forceinline__ void sdf_finish(const SdfArgs& a, int64_t i, const double q[3],
                                           const MinCand& m, bool cn) {
  double n[3];
  double T[9];
  load_tri(a.tris, m.pos, T);
  if (m.type == 2) {
    tri_normal(T, n);
  } else {
    n[0] = m.sn[0];
    n[1] = m.sn[1];
    n[2] = m.sn[2];
  }
  double sgn = 1.0;
  if (cn) {
    const bool inside_dom = !(m.p[0] < a.dom[0] || m.p[0] > a.dom[3] || m.p[1] < a.dom[1] ||
                              m.p[1] > a.dom[4] || m.p[2] < a.dom[2] || m.p[2] > a.dom[5]);
    if (!(a.watertight && !inside_dom)) {
      const double dp = dot3(q[0] - m.p[0], q[1] - m.p[1], q[2] - m.p[2], n[0], n[1], n[2]);
      sgn = (dp >= 0.0) ? 1.0 : -1.0;
    }
  }
  // a walk that took no triangle reports the initial minimum DBL_MAX (a NaN
  // query walks with bound 0, see query_has_nan)
  a.dist[i] = sqrt(m.type < 0 ? kDblMax : m.sq) * sgn;
  if (a.cp) {
    a.cp[3 * i] = m.p[0];
    a.cp[3 * i + 1] = m.p[1];
    a.cp[3 * i + 2] = m.p[2];
  }
  if (a.normal) {
    double un[3];
    unit3(n, un);
    a.normal[3 * i] = un[0];
    a.normal[3 * i + 1] = un[1];
    a.normal[3 * i + 2] = un[2];
  }
  // -1 when no triangle was taken (a query with a NaN or infinite coordinate)
  if (a.elem) a.elem[i] = m.type < 0 ? -1 : reinterpret_cast<const int32_t*>(a.tris + m.pos)[18];
}

// A walk the first pass stopped at its step budget: the whole walk state, so
// the second pass continues exactly where it stopped. The pseudo-normal sum
// is its list of terms (the sum itself is zero until the terms are added).
struct SdfSuspended {
  double sq, p[3];
  int32_t qidx, cur, sp, visits, steps, leaves, type, pos;
  int32_t nterms;
  int32_t terms[kNormList];
  int32_t stack[kStack];  // entries 1..sp
};
static_assert(sizeof(SdfSuspended) == 360, "record size quoted in the capacity comment");

// box key, pruning test and nearer-first test of the two walks
template <bool MIXED>
struct BoxTests {
  using Scalar = typename std::conditional<MIXED, float, double>::type;
  double q[3];
  float qlo[3], qhi[3], qn[3];
  double last_sq;
  float thr;
  __device__ __forceinline__ explicit BoxTests(const double qq[3]) : last_sq(-1.0), thr(0.0f) {
    for (int d = 0; d < 3; ++d) {
      q[d] = qq[d];
      qlo[d] = __double2float_rd(qq[d]);
      qhi[d] = __double2float_ru(qq[d]);
      qn[d] = (float)qq[d];
    }
  }
  // squared distance to the box; for the FP32 boxes, the sum of squared gaps
  // to the interval [rd(q), ru(q)] in round-to-nearest FP32
  __device__ __forceinline__ Scalar key(const Scalar* b) const {
    if constexpr (!MIXED) {
      return sqdist_box(q[0], q[1], q[2], b);
    } else {
      float acc = 0.0f;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const float g = fmaxf(fmaxf(b[d] - qhi[d], qlo[d] - b[3 + d]), 0.0f);
        acc += g * g;
      }
      return acc;
    }
  }
  // squared_distance(p, box) <= minSqDist; for the FP32 boxes the threshold
  // ru(minSqDist * (1 + 1e-6) + 1e-36) covers the FP32 rounding errors
  __device__ __forceinline__ bool keep(Scalar k, double min_sq) {
    if constexpr (!MIXED) {
      return k <= min_sq;
    } else {
      if (min_sq != last_sq) {
        last_sq = min_sq;
        thr = __double2float_ru(min_sq * (1.0 + 1e-6) + 1e-36);
      }
      return k <= thr;
    }
  }
  // LinearBVH.hpp:75-82: true when the right child's centroid is nearer
  __device__ __forceinline__ bool right_first(const Scalar* l, const Scalar* r) const {
    if constexpr (!MIXED) {
      const double cl0 = 0.5 * (l[0] + l[3]), cl1 = 0.5 * (l[1] + l[4]), cl2 = 0.5 * (l[2] + l[5]);
      const double cr0 = 0.5 * (r[0] + r[3]), cr1 = 0.5 * (r[1] + r[4]), cr2 = 0.5 * (r[2] + r[5]);
      const double dl = dot3(cl0 - q[0], cl1 - q[1], cl2 - q[2], cl0 - q[0], cl1 - q[1], cl2 - q[2]);
      const double dr = dot3(cr0 - q[0], cr1 - q[1], cr2 - q[2], cr0 - q[0], cr1 - q[1], cr2 - q[2]);
      return dl > dr;
    } else {
      float dl = 0.0f, dr = 0.0f;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const float el = 0.5f * (l[d] + l[3 + d]) - qn[d];
        const float er = 0.5f * (r[d] + r[3 + d]) - qn[d];
        dl += el * el;
        dr += er * er;
      }
      return dl > dr;
    }
  }
};

// First pass: every query, in Morton order, one lane each, with a step budget
// and the vote-scheduled walk. A walk over its budget is saved (while there
// is room) for the second pass; a walk whose term list overflowed is queued to
// be run again from the start there. Splitting keeps the rare very long walks
// (queries near the centre of a closed surface, where thousands of triangles
// are almost equidistant) out of the coherent waves of this pass, which would
// otherwise all wait for them. CELLS: a mesh with quad cells, whose leaf
// holds its cell's one or two triangle records at tris[2 * pos + e], checked
// in order (SignedDistance.hpp:649-658); MinCand::pos and the term list then
// name records, not leaves.
template <bool MIXED, bool STATS, bool CELLS = false>
__global__ __launch_bounds__(kSdfBlock) AXSQ_SDF_WAVES_ATTR void k_sdf_eval(SdfArgs a) {
  using NodeT = typename std::conditional<MIXED, Node32, Node64>::type;
  using Scalar = typename std::conditional<MIXED, float, double>::type;
  __shared__ int32_t s_stack[stack_lds_words(kSdfBlock)];
  __shared__ int32_t s_terms[kNormList * kSdfBlock];
  LaneStack todo;
  todo.lds = lane_stack_base(s_stack);
  lds_i32* terms =
      (lds_i32*)(s_terms + (threadIdx.x / kWaveLanes) * (kNormList * kWaveLanes) +
                 threadIdx.x % kWaveLanes);
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t t = g;  // position in the query order
  if (a.wave_order) {
    const int64_t wl = g / kWaveLanes;
    if (wl >= (a.nq + kWaveLanes - 1) / kWaveLanes) return;
    t = (int64_t)a.wave_order[wl] * kWaveLanes + g % kWaveLanes;
  }
  if (t >= a.nq) return;
  const int64_t i = a.order ? (int64_t)a.order[t] : t;
  if (a.skip && a.skip[i]) {  // walked beside this pass from the start
    if (a.wsteps) a.wsteps[t] = 0;
    return;
  }
  MinCand m;
  m.sq = kDblMax;
  m.p[0] = m.p[1] = m.p[2] = 0.0;
  m.sn[0] = m.sn[1] = m.sn[2] = 0.0;
  m.type = -1;
  m.pos = 0;
  Walk w = walk_start(todo);
  // this pass's steps of the walk, for the next call's launch order, on every
  // way out
  struct StepsOut {
    int32_t* p;
    const Walk& w;
    __device__ ~StepsOut() {
      if (p) *p = w.steps;
    }
  } steps_out{a.wsteps ? a.wsteps + t : nullptr, w};
  int32_t leaves = 0, nterms = 0;
  int32_t budget = a.budget;
  const double q[3] = {a.qx[i * a.qs], a.qy[i * a.qs], a.qz[i * a.qs]};
  if (query_has_nan(q)) m.sq = 0.0;
  const bool cn = a.compute_sign != 0;
  const NodeT* nodes;
  if constexpr (MIXED)
    nodes = a.nodes32;
  else
    nodes = a.nodes;
  BoxTests<MIXED> bt(q);
  auto pred = [&](const Scalar* b) { return bt.keep(bt.key(b), m.sq); };
  auto pref = [&](const Scalar* l, const Scalar* r) { return bt.right_first(l, r); };
  auto leaf = [&](int32_t pos) {
    double T[9];
    if constexpr (CELLS) {
      const int32_t ntri = reinterpret_cast<const int32_t*>(a.tris + 2 * (int64_t)pos)[19];
      for (int e = 0; e < 2; ++e) {
        if (e < ntri) {
          load_tri(a.tris, 2 * pos + e, T);
          check_candidate_deferred(m, q, T, 2 * pos + e, cn, terms, nterms);
        }
      }
    } else {
      load_tri(a.tris, pos, T);
      check_candidate_deferred(m, q, T, pos, cn, terms, nterms);
    }
    if (STATS) ++leaves;
  };
  int st;
  for (;;) {
#if AXSQ_SDF_VOTE
    st = traverse_vote<NodeT, Scalar>(nodes, a.ninner, todo, w, budget, pred, pref, leaf);
#else
    st = traverse<NodeT, Scalar>(nodes, a.ninner, todo, w, budget, pred, pref, leaf);
#endif
    if (st != kWalkSuspended || (nterms & kListOverflow)) break;
    const uint32_t slot = atomicAdd(a.nsusp, 1u);
    if (slot < (uint32_t)a.cap) {
      SdfSuspended& r = a.susp[slot];
      r.qidx = (int32_t)i;
      r.sq = m.sq;
      for (int k = 0; k < 3; ++k) r.p[k] = m.p[k];
      r.type = m.type;
      r.pos = m.pos;
      r.cur = w.cur;
      r.sp = w.sp;
      r.visits = w.visits;
      r.steps = w.steps;
      r.leaves = leaves;
      r.nterms = nterms;
      for (int32_t k = 0; k < nterms; ++k) r.terms[k] = terms[k * kWaveLanes];
      for (int32_t k = 1; k <= w.sp; ++k) r.stack[k] = todo.get(k);
      return;
    }
    budget = 0x7fffffff;  // no room left: finish here
  }
  if (st == kWalkError) {
    atomicOr(a.status, kStatusStack);
  } else if (nterms & kListOverflow) {
    a.redo[atomicAdd(a.nredo, 1u)] = (int32_t)i;  // one slot per query
    return;
  }
  if (STATS) {
    atomicAdd((unsigned long long*)&a.stats[0], (unsigned long long)(st == kWalkError ? 0 : w.visits));
    atomicAdd((unsigned long long*)&a.stats[1], (unsigned long long)leaves);
  }
  add_terms(m.sn, a.tris, (const lds_i32*)terms, kWaveLanes, nterms & (kListOverflow - 1));
  sdf_finish(a, i, q, m, cn);
}

// ---------------------------------------------------------------------------
// Second pass, wave-cooperative form: one walk per wave, either a saved walk
// or (after a term-list overflow) a walk run again from the start. The walk
// itself is sequential (its pruning depends on every earlier leaf), so all 64
// lanes run it in lockstep on wave-uniform state, and the lanes are put to
// work on what does not depend on that state: when the walk enters a subtree
// of at most 64 leaves, lane k computes, for this query, the closest point and
// normal term of leaf lo + k and the two child-box keys and the nearer-first
// bit of inner node nb + k, in registers. The walk then reads them with
// v_readlane instead of loading a node or a triangle and running the closest
// point on the critical path. Each memo entry is computed by the same
// operations as in the one-lane walk, and the walk visits nodes and leaves in
// the same order, so the results are bit-identical. The stack is held one
// entry per lane (entry k in lane k). Normal terms are added as they come.
constexpr double kSdfEps = 1e-12;  // SignedDistance.hpp:647 (closest point, normal ties)

// checkCandidate's per-triangle values (SignedDistance.hpp:637-737): squared
// distance, closest point, location type and the code of the pseudo-normal
// term the triangle adds if it ties an edge (3) or vertex (0..2) closest
// point. The term itself is only computed when it is added (rarely).
__device__ __forceinline__ void leaf_record(const double q[3], const double T[9], double& sq,
                                            double cp[3], int32_t& type, int32_t& code) {
  const int loc = closest_point_ilp(q, T, kSdfEps, cp);
  sq = dot3(cp[0] - q[0], cp[1] - q[1], cp[2] - q[2], cp[0] - q[0], cp[1] - q[1], cp[2] - q[2]);
  type = (loc < 0) ? 1 : (loc < 3) ? 0 : 2;
  code = (loc < 0) ? 3 : loc;
}

// The state update of checkCandidate given leaf_record's values.
__device__ __forceinline__ void apply_record(MinCand& m, const TriRec* __restrict__ tris,
                                             int32_t pos, double sq, const double cp[3],
                                             int32_t type, int32_t code, bool cn) {
  const double e0 = m.p[0] - cp[0], e1 = m.p[1] - cp[1], e2 = m.p[2] - cp[2];
  const bool same = near_eq(dot3(e0, e1, e2, e0, e1, e2), 0., kSdfEps);
  bool upd;
  if (sq < m.sq) {
    const bool clear = (type == 2) || (m.type != type) || !same;
    m.sq = sq;
    m.p[0] = cp[0];
    m.p[1] = cp[1];
    m.p[2] = cp[2];
    m.type = type;
    m.pos = pos;
    if (cn && clear) m.sn[0] = m.sn[1] = m.sn[2] = 0.0;
    upd = cn && (type != 2);
  } else {
    upd = cn && (type != 2) && (m.type == type) && same;
  }
  if (upd) {
    double T[9], c[3];
    load_tri(tris, pos, T);
    normal_term(T, code, c);
    m.sn[0] += c[0];
    m.sn[1] += c[1];
    m.sn[2] += c[2];
  }
}

__device__ __forceinline__ int32_t lane_get(int32_t v, int32_t l) {
  return __builtin_amdgcn_readlane(v, l);
}
__device__ __forceinline__ float lane_get(float v, int32_t l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_get(double v, int32_t l) {
  const int64_t b = __double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int32_t)b, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int32_t)(b >> 32), l);
  return __longlong_as_double((int64_t)(((uint64_t)hi << 32) | lo));
}

// One lane's share of a subtree memo: leaf lo + lane and inner node nb + lane.
template <class Scalar>
struct MemoLane {
  double sq, p[3];
  int32_t type, code;
  double sq2, p2[3];      // quad cells: the cell's second triangle (type2 -2: none)
  int32_t type2, code2;
  Scalar lk, rk;  // child-box keys (BoxTests::key)
  int32_t lc, rc;
  int32_t bits;     // bit 0/1: child box valid, bit 2: right child first
  int32_t nlo, nhi; // the node's leaf range
  int32_t pn;       // this node's parent in the memo, 2 * index + (1 if right child); -1: root
  int32_t pl;       // this leaf's parent in the memo, 2 * index + (1 if right child)
};

// A hot walk's tables: for its query, every leaf's record (leaf_record's
// values) and every inner node's child-box keys and bits (MemoLane::bits),
// computed by the same functions as a memo fill, so a fill that reads them
// holds the same values.
struct HotLeaf {
  double sq, p[3];
  int32_t type, code;
};
template <class Scalar>
struct HotNode {
  Scalar lk, rk;
  int32_t bits;
};

// bits [lo, hi) of a 64-bit mask (0 <= lo <= hi <= 64)
__device__ __forceinline__ uint64_t bit_range(int32_t lo, int32_t hi) {
  const uint64_t below_hi = hi >= 64 ? ~uint64_t(0) : ((uint64_t(1) << hi) - 1);
  const uint64_t below_lo = lo >= 64 ? ~uint64_t(0) : ((uint64_t(1) << lo) - 1);
  return below_hi & ~below_lo;
}

// CELLS: quad cells, each leaf holding one or two triangle records (as in
// k_sdf_eval); a memo lane then holds both records of its leaf, and a leaf
// changes the running state when either record would.
template <bool MIXED, bool STATS, bool CELLS = false>
__global__ __launch_bounds__(kWaveLanes) AXSQ_RESUME_WAVES_ATTR void k_sdf_resume_wave(SdfArgs a, int64_t nres,
                                                                 int64_t nredo) {
  static_assert(kStack == kWaveLanes, "one stack entry per lane");
  using NodeT = typename std::conditional<MIXED, Node32, Node64>::type;
  using Scalar = typename std::conditional<MIXED, float, double>::type;
  __shared__ int32_t s_pn[kWaveLanes], s_pl[kWaveLanes], s_seq[2 * kWaveLanes];
  const int32_t lane = (int32_t)threadIdx.x;
  const int64_t b = blockIdx.x;
  if (b >= nres + nredo) return;
  const bool cn = a.compute_sign != 0;
  int64_t i;
  MinCand m;
  int32_t cur, sp, visits, leaves;
  // the stack, entry k in lane k, with the leaf range of each node entry
  // (lo < 0: not known, read from the node when needed)
  int32_t stk, stk_lo = -1, stk_hi = -1;
  int32_t cur_lo = -1, cur_hi = -1;  // leaf range of cur, if known
  if (b < nres) {
    const SdfSuspended& r = a.susp[a.susp_order ? (int64_t)a.susp_order[b] : b];
    i = r.qidx;
    m.sq = r.sq;
    for (int k = 0; k < 3; ++k) m.p[k] = r.p[k];
    m.type = r.type;
    m.pos = r.pos;
    cur = r.cur;
    sp = r.sp;
    visits = r.visits;
    leaves = r.leaves;
    if (sp < 0 || sp >= kStack || r.nterms < 0 || r.nterms > kNormList) {
      if (lane == 0) atomicOr(a.status, kStatusStack);
      return;
    }
    stk = (lane == 0) ? kBarrier : (lane <= sp ? r.stack[lane] : 0);
    m.sn[0] = m.sn[1] = m.sn[2] = 0.0;
    add_terms(m.sn, a.tris, r.terms, 1, r.nterms);
  } else {
    i = a.redo[b - nres];
    m.sq = kDblMax;
    m.p[0] = m.p[1] = m.p[2] = 0.0;
    m.sn[0] = m.sn[1] = m.sn[2] = 0.0;
    m.type = -1;
    m.pos = 0;
    cur = 0;
    cur_lo = 0;
    cur_hi = a.ninner;
    sp = 0;
    visits = 0;
    leaves = 0;
    stk = kBarrier;
  }
  i = (i >= 0 && i < a.nq) ? i : 0;  // only a malformed record or list is out of range
  const double q[3] = {a.qx[i * a.qs], a.qy[i * a.qs], a.qz[i * a.qs]};
  if (b >= nres && query_has_nan(q)) m.sq = 0.0;  // a walk from the start
  const NodeT* nodes;
  if constexpr (MIXED)
    nodes = a.nodes32;
  else
    nodes = a.nodes;
  BoxTests<MIXED> bt(q);
  const HotLeaf* hl = nullptr;
  const HotNode<Scalar>* hn = nullptr;
  if (b < a.nhot) {
    hl = static_cast<const HotLeaf*>(a.hot_leaf) + b * (int64_t)(a.ninner + 1);
    hn = static_cast<const HotNode<Scalar>*>(a.hot_node) + b * (int64_t)a.ninner;
  }

  auto pop = [&]() {
    cur = lane_get(stk, sp);
    cur_lo = lane_get(stk_lo, sp);
    cur_hi = lane_get(stk_hi, sp);
    sp--;
  };
  auto push = [&](int32_t v, int32_t vlo, int32_t vhi) {  // sp already raised
    stk = (lane == sp) ? v : stk;
    stk_lo = (lane == sp) ? vlo : stk_lo;
    stk_hi = (lane == sp) ? vhi : stk_hi;
  };

  MemoLane<Scalar> mem;
  int32_t clo = 1, chi = 0;  // memo leaves [clo, chi] (empty)
  int32_t nb = 0, nn = 0;    // memo inner nodes [nb, nb + nn)
  // a leaf held back by the walk when a new memo replaces the old one (a
  // memo entry depends only on the query and the leaf, so it stays valid)
  int32_t h_pos = -1, h_type = 0;
  int32_t h_code = 0;
  double h_sq = 0.0, h_p[3] = {0.0, 0.0, 0.0};
  int32_t h_type2 = -2, h_code2 = 0;  // quad cells: the held leaf's second record
  double h_sq2 = 0.0, h_p2[3] = {0.0, 0.0, 0.0};
  // whether a record (sq, cp, type) would change the running state: a strictly
  // closer point, or a pseudo-normal term to add (apply_record's two cases)
  auto rec_eff = [&](double rsq, const double rp[3], int32_t rtype) -> bool {
    const double e0 = m.p[0] - rp[0], e1 = m.p[1] - rp[1], e2 = m.p[2] - rp[2];
    return (rsq < m.sq) ||
           (cn && rtype != 2 && m.type == rtype && near_eq(dot3(e0, e1, e2, e0, e1, e2), 0., kSdfEps));
  };
  // a quad cell's records from the geometry (the second: type -2 when absent)
  auto cell_records = [&](int32_t pos, double& s1, double p1[3], int32_t& t1, int32_t& c1,
                          double& s2, double p2[3], int32_t& t2, int32_t& c2) {
    double T[9];
    const int32_t ntri = reinterpret_cast<const int32_t*>(a.tris + 2 * (int64_t)pos)[19];
    load_tri(a.tris, 2 * pos, T);
    leaf_record(q, T, s1, p1, t1, c1);
    if (ntri >= 2) {
      load_tri(a.tris, 2 * pos + 1, T);
      leaf_record(q, T, s2, p2, t2, c2);
    } else {
      s2 = kDblMax;
      p2[0] = p2[1] = p2[2] = 0.0;
      t2 = -2;
      c2 = 0;
    }
  };

  // Wave masks over the memo, bit k for lane k: the node's left / right child
  // passes the pruning test at the current minimum, the node visits its right
  // child first, the leaf would change the running state (a strictly closer
  // point, or a pseudo-normal term to add). The minimum only changes at a leaf
  // of the last kind, after which the masks are recomputed, so a node visit is
  // a few bit tests and a leaf without effect costs nothing.
  uint64_t m_l = 0, m_r = 0, m_sw = 0, m_eff = 0;
  auto remask = [&]() {
    const bool node_ok = lane < nn;
    m_l = __ballot(node_ok && (mem.bits & 1) && bt.keep(mem.lk, m.sq));
    m_r = __ballot(node_ok && (mem.bits & 2) && bt.keep(mem.rk, m.sq));
    bool eff = false;
    if (lane <= chi - clo) {
      if constexpr (CELLS) {
        eff = rec_eff(mem.sq, mem.p, mem.type) || rec_eff(mem.sq2, mem.p2, mem.type2);
      } else {
        const double e0 = m.p[0] - mem.p[0], e1 = m.p[1] - mem.p[1], e2 = m.p[2] - mem.p[2];
        eff = (mem.sq < m.sq) ||
              (cn && mem.type != 2 && m.type == mem.type &&
               near_eq(dot3(e0, e1, e2, e0, e1, e2), 0., kSdfEps));
      }
    }
    m_eff = __ballot(eff);
  };
  auto fill = [&](int32_t c, int32_t lo, int32_t hi) {
    const int32_t nb_new = (c == lo) ? lo : lo + 1;
    const int32_t nn_new = hi - lo;
    // both loads first (indices clamped into the subtree, so every lane
    // loads), so that their latencies overlap; lanes past the end keep
    // values nothing reads
    Scalar lb[6], rb[6];
    int32_t flags;
    const int32_t my_node = nb_new + min(lane, nn_new - 1);
    const int32_t my_leaf = lo + min(lane, hi - lo);
    if constexpr (CELLS) {
      load_node(nodes, my_node, lb, rb, mem.lc, mem.rc, flags);
      cell_records(my_leaf, mem.sq, mem.p, mem.type, mem.code, mem.sq2, mem.p2, mem.type2,
                   mem.code2);
      mem.lk = bt.key(lb);
      mem.rk = bt.key(rb);
      mem.bits = (flags & 3) | (bt.right_first(lb, rb) ? 4 : 0);
    } else if (hl) {  // a hot walk: the tables hold the leaf records and node keys
      const HotLeaf r = hl[my_leaf];
      const HotNode<Scalar> k = hn[my_node];
      load_node(nodes, my_node, lb, rb, mem.lc, mem.rc, flags);  // links only
      mem.sq = r.sq;
      for (int d = 0; d < 3; ++d) mem.p[d] = r.p[d];
      mem.type = r.type;
      mem.code = r.code;
      mem.lk = k.lk;
      mem.rk = k.rk;
      mem.bits = k.bits;
    } else {
      double T[9];
      load_tri(a.tris, my_leaf, T);
      load_node(nodes, my_node, lb, rb, mem.lc, mem.rc, flags);
      leaf_record(q, T, mem.sq, mem.p, mem.type, mem.code);
      mem.lk = bt.key(lb);
      mem.rk = bt.key(rb);
      mem.bits = (flags & 3) | (bt.right_first(lb, rb) ? 4 : 0);
    }
    int2 rg;
    if constexpr (MIXED) {
      const int4 r4 = a.ranges32[my_node];
      rg = make_int2(r4.x, r4.y);
    } else {
      rg = *reinterpret_cast<const int2*>(&a.nodes[my_node].lo);
    }
    mem.nlo = rg.x;
    mem.nhi = rg.y;
    // parent links inside the memo, by scatter through LDS
    s_pn[lane] = -1;
    s_pl[lane] = -1;
    __syncthreads();
    if (lane < nn_new) {
      const int32_t ch[2] = {mem.lc, mem.rc};
      for (int side = 0; side < 2; ++side) {
        const int32_t c2 = ch[side];
        if (c2 >= 0) {
          if (c2 - nb_new >= 0 && c2 - nb_new < nn_new) s_pn[c2 - nb_new] = 2 * lane + side;
        } else {
          const int32_t j = (-c2 - 1) - lo;
          if (j >= 0 && j <= hi - lo) s_pl[j] = 2 * lane + side;
        }
      }
    }
    __syncthreads();
    mem.pn = s_pn[lane];
    mem.pl = s_pl[lane];
    clo = lo;
    chi = hi;
    nb = nb_new;
    nn = nn_new;
    m_sw = __ballot(lane < nn && (mem.bits & 4));
    remask();
  };
  auto leaf = [&](int32_t pos) {
    double sq, cp[3];
    int32_t type, code;
    if constexpr (CELLS) {  // the cell's records in order, from the memo or the geometry
      double sq2, cp2[3];
      int32_t type2, code2;
      if (pos >= clo && pos <= chi) {
        const int32_t k = pos - clo;
        if (!((m_eff >> k) & 1)) return;  // no change to the running state
        sq = lane_get(mem.sq, k);
        sq2 = lane_get(mem.sq2, k);
        for (int d = 0; d < 3; ++d) {
          cp[d] = lane_get(mem.p[d], k);
          cp2[d] = lane_get(mem.p2[d], k);
        }
        type = lane_get(mem.type, k);
        code = lane_get(mem.code, k);
        type2 = lane_get(mem.type2, k);
        code2 = lane_get(mem.code2, k);
      } else if (pos == h_pos) {
        sq = h_sq;
        sq2 = h_sq2;
        for (int d = 0; d < 3; ++d) {
          cp[d] = h_p[d];
          cp2[d] = h_p2[d];
        }
        type = h_type;
        code = h_code;
        type2 = h_type2;
        code2 = h_code2;
      } else {
        cell_records(pos, sq, cp, type, code, sq2, cp2, type2, code2);
      }
      apply_record(m, a.tris, 2 * pos, sq, cp, type, code, cn);
      if (type2 != -2) apply_record(m, a.tris, 2 * pos + 1, sq2, cp2, type2, code2, cn);
      remask();
      return;
    }
    if (pos >= clo && pos <= chi) {
      const int32_t k = pos - clo;
      if (!((m_eff >> k) & 1)) return;  // no change to the running state
      sq = lane_get(mem.sq, k);
      for (int d = 0; d < 3; ++d) cp[d] = lane_get(mem.p[d], k);
      type = lane_get(mem.type, k);
      code = lane_get(mem.code, k);
    } else if (pos == h_pos) {
      sq = h_sq;
      for (int d = 0; d < 3; ++d) cp[d] = h_p[d];
      type = h_type;
      code = h_code;
    } else if (hl) {
      const HotLeaf r = hl[pos];
      sq = r.sq;
      for (int d = 0; d < 3; ++d) cp[d] = r.p[d];
      type = r.type;
      code = r.code;
    } else {
      double T[9];
      load_tri(a.tris, pos, T);
      leaf_record(q, T, sq, cp, type, code);
    }
    apply_record(m, a.tris, pos, sq, cp, type, code, cn);
    remask();
  };

  // The walk of traverse() (bvh_traverse.hpp:66-154), resumed, written as a
  // state machine: kInner -- cur is the next inner node to visit and `found`
  // the held leaf (or 0); kLeaf -- the leaf loop, `found` is the next leaf to
  // check. A tight loop takes every step it can from the memo masks alone
  // (scalar code, no memory access); anything else -- a node outside the
  // memo, a leaf that changes the running state, a guard -- is one step of
  // the general path, after which the tight loop resumes.
  enum { kInner = 0, kLeaf = 1, kDone = 2 };
  int phase = kInner;  // a saved walk stops at the top of its outer loop
  int32_t found = 0;
  bool err = false;
  if (cur == kBarrier) phase = kDone;
  else if (cur < 0) err = true;  // only from a malformed record
  // after a node visit: hold a first leaf, leave the inner loop at a second
  auto after_visit = [&]() {
    if (cur < 0 && !(found < 0)) {
      found = cur;
      if (cur != kBarrier) {
        if (sp < 0) {
          err = true;
          return;
        }
        pop();
      }
    }
    if (cur < 0) {
      if (found < 0 && found != kBarrier) {
        phase = kLeaf;
      } else {
        found = 0;
        phase = (cur == kBarrier) ? kDone : kInner;
        if (phase == kInner) err = true;  // cur < 0 here only for a malformed tree
      }
    }
  };
  // after a leaf check: next leaf, or back to the top of the outer loop
  auto after_leaf = [&]() {
    found = cur;
    if (cur < 0 && cur != kBarrier) {
      if (sp < 0) {
        err = true;
        return;
      }
      pop();
    }
    if (!(found < 0 && found != kBarrier)) {
      found = 0;
      phase = (cur == kBarrier) ? kDone : kInner;
    }
  };
  // A subtree whose walk changes nothing: the walk has just reached the
  // memo's root (cur), and no leaf it would check changes the running state
  // (nor the held leaf, if any, which it would check at the first leaf). The
  // minimum then stays fixed through the subtree, so the nodes it visits and
  // the leaves it reaches follow from the memo masks alone, and so does their
  // order (pre-order, nearer child first). They are found for all memo nodes
  // at once: by pointer jumping along parent links (a node is visited when no
  // link on its path is cut; its position is the sum along its path of its
  // offsets after its parent), and the item count under a node is a
  // popcount over its contiguous node and leaf ranges. What the walk state
  // needs afterwards is the visit and check counts and which leaf, if any, is
  // held when the walk leaves the subtree; the latter follows from the runs
  // of consecutive leaves in that order (a lone leaf is held, or frees a held
  // one; a run of two or more is checked in the leaf loop). Returns false,
  // changing nothing, when the case does not apply.
  auto skip_subtree = [&]() -> bool {
    const int32_t r = cur - nb;
    if (r < 0 || r >= nn || sp < 0) return false;
    const bool held = (found < 0 && found != kBarrier);
    bool held_eff = true;
    if (held && -found - 1 == h_pos) {
      if constexpr (CELLS) {
        held_eff = rec_eff(h_sq, h_p, h_type) || rec_eff(h_sq2, h_p2, h_type2);
      } else {
        const double e0 = m.p[0] - h_p[0], e1 = m.p[1] - h_p[1], e2 = m.p[2] - h_p[2];
        held_eff = (h_sq < m.sq) || (cn && h_type != 2 && m.type == h_type &&
                                     near_eq(dot3(e0, e1, e2, e0, e1, e2), 0., kSdfEps));
      }
    }
    const bool node_lane = lane < nn;
    const bool leaf_lane = lane <= chi - clo;
    // parent lane (itself for the root and for lanes outside the memo) and
    // whether the parent lets this node through
    const int32_t pn = (node_lane && lane != r) ? mem.pn : -1;
    const int32_t pnode = pn >= 0 ? (pn >> 1) : lane;
    const bool cut = pn >= 0 && !((((pn & 1) ? m_r : m_l) >> pnode) & 1u);
    // visited: no cut on the path to the root
    int32_t anc = pnode, acc = cut ? 1 : 0;
    int32_t rounds = 0;
    for (;;) {
      const int32_t acc_a = __shfl(acc, anc), anc_a = __shfl(anc, anc);
      acc |= acc_a;
      anc = anc_a;
      ++rounds;
      if (__ballot(node_lane && anc != r && anc != lane) == 0 || rounds > 7) break;
    }
    if (rounds > 7) return false;
    const uint64_t vis = __ballot(node_lane && acc == 0);
    if (!((vis >> r) & 1u)) return false;
    bool lr = false;
    if (leaf_lane && mem.pl >= 0) {
      const int32_t p = mem.pl >> 1;
      const uint64_t pass = (mem.pl & 1) ? m_r : m_l;
      lr = ((vis >> p) & 1u) && ((pass >> p) & 1u);
    }
    const uint64_t reached = __ballot(leaf_lane && lr);
    if (reached & m_eff) return false;
    const int32_t n_nodes = __popcll(vis), n_leaves = __popcll(reached);
    if (held && n_leaves > 0 && held_eff) return false;
    if (visits + n_nodes > a.ninner) return false;  // the visit guard would trip
    // items under each visited node: popcounts over its node and leaf ranges
    const bool mine = node_lane && ((vis >> lane) & 1u);
    int32_t size = 0;
    if (mine) {
      const int32_t first_node = ((nb + lane) == mem.nlo ? mem.nlo : mem.nlo + 1) - nb;
      const int32_t n_in = mem.nhi - mem.nlo;
      size = __popcll(vis & bit_range(first_node, first_node + n_in)) +
             __popcll(reached & bit_range(mem.nlo - clo, mem.nhi - clo + 1));
    }
    const bool l_pass = mine && ((m_l >> lane) & 1u), r_pass = mine && ((m_r >> lane) & 1u);
    const bool l_inner = mem.lc >= 0, r_inner = mem.rc >= 0;
    const int32_t li = (node_lane && l_inner) ? mem.lc - nb : lane;
    const int32_t ri = (node_lane && r_inner) ? mem.rc - nb : lane;
    const int32_t sl = __shfl(size, li), sr = __shfl(size, ri);
    const int32_t size_l = l_pass ? (l_inner ? sl : 1) : 0;
    const int32_t size_r = r_pass ? (r_inner ? sr : 1) : 0;
    // pre-order: the first child follows its parent, the second follows the
    // first child's items
    const bool right_first = (l_pass && r_pass) ? (((m_sw >> lane) & 1u) != 0) : !l_pass;
    const int32_t first_size = right_first ? size_r : size_l;
    const int32_t both = (l_pass && r_pass) ? 1 : 0;
    auto offset = [&](int32_t link, int32_t p) -> int32_t {
      const int32_t p_rf = __shfl((int32_t)right_first, p);
      const int32_t p_both = __shfl(both, p);
      const int32_t p_fs = __shfl(first_size, p);
      const bool second = p_both && ((link & 1) == (p_rf ? 0 : 1));
      return 1 + (second ? p_fs : 0);
    };
    const int32_t node_off = offset(pn, pnode);
    const int32_t pleaf = (leaf_lane && mem.pl >= 0) ? (mem.pl >> 1) : lane;
    const int32_t leaf_off = offset(leaf_lane ? mem.pl : 0, pleaf);
    // positions and depths: sums along the path, by pointer jumping
    int32_t pre = (mine && lane != r) ? node_off : 0;
    int32_t dep = (mine && lane != r) ? 1 : 0;
    anc = (mine && lane != r) ? pnode : lane;
    for (int32_t k = 0; k < rounds; ++k) {
      const int32_t pre_a = __shfl(pre, anc), dep_a = __shfl(dep, anc), anc_a = __shfl(anc, anc);
      pre += pre_a;
      dep += dep_a;
      anc = anc_a;
    }
    // a node whose children both pass pushes one entry above the at most
    // `dep` its ancestors left: leave an overflow to the step-by-step walk
    if (__ballot(mine && sp + dep + 1 >= kStack) != 0) return false;
    const int32_t lpos = __shfl(pre, pleaf) + leaf_off;
    const int32_t total = lane_get(size, r);
    if (total != n_nodes + n_leaves || total > 2 * kWaveLanes) return false;
    // leaf at each position (memo leaf index), -1 for a node
    s_seq[lane] = -1;
    s_seq[kWaveLanes + lane] = -1;
    __syncthreads();
    if (leaf_lane && lr && lpos >= 0 && lpos < total) s_seq[lpos] = lane;
    __syncthreads();
    const int32_t q0 = s_seq[lane], q1 = s_seq[kWaveLanes + lane];
    const uint64_t lm0 = __ballot(q0 >= 0), lm1 = __ballot(q1 >= 0);
    if (__popcll(lm0) + __popcll(lm1) != n_leaves) return false;
    // Runs of consecutive leaves in that order. Walk states: 0 nothing held,
    // 1 one leaf held, 2 in the leaf loop. A lone leaf met with nothing held
    // is held; a lone leaf met while one is held checks both; a run of two or
    // more checks all of them (and a held one); a node visit ends the leaf
    // loop. So every reached leaf is checked except one held at the exit, and
    // the exit state follows from the last run: after a long run (or a lone
    // leaf that found one held) the walk is in the leaf loop, or at state 0
    // if nodes follow; a lone last leaf is held when the lone runs since the
    // last long run (or since the start, from the entry state) leave nothing
    // held before it.
    int32_t state, held_pos = held ? (-found - 1) : -1;
    int32_t checks = n_leaves + ((held && n_leaves > 0) ? 1 : 0);
    if (n_leaves == 0) {
      state = held ? 1 : 0;
    } else {
      typedef unsigned __int128 u128;
      const u128 lmask = ((u128)lm1 << 64) | (u128)lm0;
      const u128 lone = lmask & ~(lmask << 1) & ~(lmask >> 1);
      const u128 longb = lmask & ~lone;
      auto top_bit = [](u128 x) -> int32_t {
        const uint64_t hi = (uint64_t)(x >> 64), lo = (uint64_t)x;
        return hi ? 127 - __clzll(hi) : (lo ? 63 - __clzll(lo) : -1);
      };
      auto count = [](u128 x) -> int32_t {
        return __popcll((uint64_t)(x >> 64)) + __popcll((uint64_t)x);
      };
      const int32_t h = top_bit(lmask);
      bool hold_at_exit = false;
      if ((lone >> h) & 1) {
        const int32_t e = top_bit(longb);  // end of the last long run, -1 if none
        const u128 between = (h > 0 ? (((u128)1 << h) - 1) : 0) & ~((((u128)1 << (e + 1)) - 1));
        const int32_t lone_before = count(lone & between);
        const int32_t base = (e >= 0) ? 0 : (held ? 1 : 0);
        const int32_t entry_h = base ^ (lone_before & 1);
        hold_at_exit = (entry_h == 0);
      }
      if (hold_at_exit) {
        state = 1;
        held_pos = clo + (h < kWaveLanes ? lane_get(q0, h) : lane_get(q1, h - kWaveLanes));
        checks -= 1;
      } else {
        state = (h < total - 1) ? 0 : 2;
      }
    }
    visits += n_nodes;
    leaves += checks;
    found = (state == 1) ? -(held_pos + 1) : 0;
    pop();
    if (state == 2) {
      phase = kLeaf;
      after_leaf();
    } else {
      phase = kInner;
      after_visit();
    }
    return true;
  };
  while (!err && phase != kDone) {
    // ---- tight loop over the memo
    for (;;) {
      if (phase == kInner) {
        const int32_t k = cur - nb;
        if ((uint32_t)k >= (uint32_t)nn || visits >= a.ninner) break;
        const uint32_t bl = (uint32_t)(m_l >> k) & 1u;
        const uint32_t br = (uint32_t)(m_r >> k) & 1u;
        if ((bl & br) && sp + 1 >= kStack) break;  // the general path reports it
        ++visits;
        if ((bl | br) == 0) {
          pop();
        } else {
          const int32_t lc = lane_get(mem.lc, k), rc = lane_get(mem.rc, k);
          if (bl & br) {
            const bool right_first = ((m_sw >> k) & 1u) != 0;
            const int32_t near = right_first ? rc : lc, far = right_first ? lc : rc;
            ++sp;
            push(far, -1, -1);  // inside the memo: no range needed
            cur = near;
          } else {
            cur = bl ? lc : rc;
          }
          cur_lo = cur_hi = -1;
        }
        after_visit();
        if (err) break;
      } else if (phase == kLeaf) {
        const int32_t k = -found - 1 - clo;
        if ((uint32_t)k > (uint32_t)(chi - clo) || ((m_eff >> k) & 1u)) break;
        ++leaves;  // a leaf that changes nothing
        after_leaf();
        if (err) break;
      } else {
        break;
      }
    }
    if (err || phase == kDone) break;
    // ---- one step of the general path
    if (phase == kInner) {
      if (cur >= a.ninner || visits >= a.ninner) {
        err = true;
        break;
      }
      const bool in_memo = cur >= nb && cur < nb + nn;
      if (!in_memo) {
        if (cur_lo < 0) {
          if constexpr (MIXED) {
            const int4 r4 = a.ranges32[cur];
            cur_lo = r4.x;
            cur_hi = r4.y;
          } else {
            const int2 rg = *reinterpret_cast<const int2*>(&a.nodes[cur].lo);
            cur_lo = rg.x;
            cur_hi = rg.y;
          }
        }
        if (cur_lo >= 0 && cur_hi <= a.ninner && cur_lo < cur_hi && cur_hi - cur_lo < kWaveLanes &&
            (cur == cur_lo || cur == cur_hi)) {
          if (found < 0 && found != kBarrier) {
            const int32_t hp = -found - 1;
            if (hp >= clo && hp <= chi) {
              const int32_t k = hp - clo;
              h_pos = hp;
              h_sq = lane_get(mem.sq, k);
              for (int d = 0; d < 3; ++d) h_p[d] = lane_get(mem.p[d], k);
              h_type = lane_get(mem.type, k);
              h_code = lane_get(mem.code, k);
              if constexpr (CELLS) {
                h_sq2 = lane_get(mem.sq2, k);
                for (int d = 0; d < 3; ++d) h_p2[d] = lane_get(mem.p2[d], k);
                h_type2 = lane_get(mem.type2, k);
                h_code2 = lane_get(mem.code2, k);
              }
            }
          }
          fill(cur, cur_lo, cur_hi);
          skip_subtree();
          continue;  // the tight loop takes it from here
        }
      }
      bool in_l, in_r, swap;
      int32_t lc, rc;
      int32_t l_lo = -1, l_hi = -1, r_lo = -1, r_hi = -1;  // children's leaf ranges
      if (in_memo) {
        const int32_t k = cur - nb;
        lc = lane_get(mem.lc, k);
        rc = lane_get(mem.rc, k);
        in_l = (m_l >> k) & 1;
        in_r = (m_r >> k) & 1;
        swap = (m_sw >> k) & 1;
      } else {
        Scalar lb[6], rb[6];
        int32_t flags;
        load_node(nodes, cur, lb, rb, lc, rc, flags);
        in_l = (flags & 1) ? bt.keep(bt.key(lb), m.sq) : false;
        in_r = (flags & 2) ? bt.keep(bt.key(rb), m.sq) : false;
        swap = bt.right_first(lb, rb);
        int32_t mid;
        if constexpr (MIXED)
          mid = a.ranges32[cur].z;
        else
          mid = a.nodes[cur].mid;
        l_lo = cur_lo;
        l_hi = mid;
        r_lo = mid + 1;
        r_hi = cur_hi;
      }
      ++visits;
      if (!in_l && !in_r) {
        pop();
      } else {
        cur = in_l ? lc : rc;
        cur_lo = in_l ? l_lo : r_lo;
        cur_hi = in_l ? l_hi : r_hi;
        if (in_l && in_r) {
          int32_t far = rc, far_lo = r_lo, far_hi = r_hi;
          if (swap) {
            far = cur;
            far_lo = cur_lo;
            far_hi = cur_hi;
            cur = rc;
            cur_lo = r_lo;
            cur_hi = r_hi;
          }
          sp++;
          if (sp >= kStack) {
            err = true;
            break;
          }
          push(far, far_lo, far_hi);
        }
      }
      after_visit();
      if (sp < -1) err = true;
    } else {
      const int32_t pos = -found - 1;
      if (pos > a.ninner) {
        err = true;
        break;
      }
      leaf(pos);
      ++leaves;
      after_leaf();
    }
  }
  if (lane != 0) return;
  if (err) atomicOr(a.status, kStatusStack);
  if (STATS) {
    atomicAdd((unsigned long long*)&a.stats[0], (unsigned long long)(err ? 0 : visits));
    atomicAdd((unsigned long long*)&a.stats[1], (unsigned long long)leaves);
  }
  sdf_finish(a, i, q, m, cn);
}


// The tables of the hot walks: item k < ninner + 1 is leaf k's record for
// the walk's query, item ninner + 1 + j inner node j's child-box keys and
// bits (grid y: the walk).
template <bool MIXED>
__global__ __launch_bounds__(kBlock) void k_hot_tables(SdfArgs a) {
  using NodeT = typename std::conditional<MIXED, Node32, Node64>::type;
  using Scalar = typename std::conditional<MIXED, float, double>::type;
  const int64_t h = blockIdx.y;
  const int64_t nleaf = (int64_t)a.ninner + 1;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= a.nhot || k >= nleaf + a.ninner) return;
  const SdfSuspended& r = a.susp[a.susp_order ? (int64_t)a.susp_order[h] : h];
  const int64_t i = r.qidx;
  const double q[3] = {a.qx[i * a.qs], a.qy[i * a.qs], a.qz[i * a.qs]};
  if (k < nleaf) {
    double T[9];
    HotLeaf out;
    load_tri(a.tris, (int32_t)k, T);
    leaf_record(q, T, out.sq, out.p, out.type, out.code);
    static_cast<HotLeaf*>(const_cast<void*>(a.hot_leaf))[h * nleaf + k] = out;
  } else {
    const int32_t j = (int32_t)(k - nleaf);
    const NodeT* nodes;
    if constexpr (MIXED)
      nodes = a.nodes32;
    else
      nodes = a.nodes;
    Scalar lb[6], rb[6];
    int32_t lc, rc, flags;
    load_node(nodes, j, lb, rb, lc, rc, flags);
    BoxTests<MIXED> bt(q);
    HotNode<Scalar> out;
    out.lk = bt.key(lb);
    out.rk = bt.key(rb);
    out.bits = (flags & 3) | (bt.right_first(lb, rb) ? 4 : 0);
    static_cast<HotNode<Scalar>*>(const_cast<void*>(a.hot_node))[h * (int64_t)a.ninner + j] = out;
  }
}

// Launch order of the saved walks: the farthest from the surface first (the
// running minimum at suspension, descending). Those are the walks that
// still have most of the tree to cover -- queries deep inside a closed
// surface -- so the second pass starts them first and fills the rest around
// them. The top 32 bits of a non-negative double order like the double.
__global__ __launch_bounds__(kBlock) void k_susp_keys(const SdfSuspended* __restrict__ susp,
                                                      int64_t n, uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ vals) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  keys[j] = ~(uint32_t)((uint64_t)__double_as_longlong(susp[j].sq) >> 32);
  vals[j] = (uint32_t)j;
}

// Morton code of each query in the mesh's bounding box (ordering only).
__global__ __launch_bounds__(kBlock) void k_query_codes(const double* __restrict__ qx,
                                                         const double* __restrict__ qy,
                                                         const double* __restrict__ qz,
                                                         int64_t qs, int64_t nq, double lo0,
                                                         double lo1, double lo2, double s0,
                                                         double s1, double s2,
                                                         uint32_t* __restrict__ keys,
                                                         uint32_t* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const double c[3] = {qx[i * qs], qy[i * qs], qz[i * qs]};
  const double minc[3] = {lo0, lo1, lo2};
  const double inv[3] = {s0, s1, s2};
  keys[i] = morton_of(c, minc, inv);
  vals[i] = (uint32_t)i;
}

// Cost of each wave of a first pass (64 consecutive positions of the query
// order): the most steps one of its walks took. One wave per wave.
__global__ __launch_bounds__(kBlock) void k_wave_cost(const int32_t* __restrict__ steps,
                                                       int64_t nq, int64_t nw,
                                                       int32_t* __restrict__ cost) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t w = g / kWaveLanes;
  if (w >= nw) return;  // whole waves
  int32_t c = (g < nq) ? steps[g] : 0;
  for (int o = kWaveLanes / 2; o > 0; o >>= 1) c = max(c, __shfl_xor(c, o));
  if (g % kWaveLanes == 0) cost[w] = c;
}

// Launch order of the first pass's waves: those whose cost falls in the
// lowest `tail_pct` percent (whole bins of a 1,024-bin histogram over
// [0, cmax], so at most that share) go last, the others first, each group in
// query order. One block of 1,024 threads: the histogram and its scan give
// the threshold, then rounds of 1,024 consecutive waves place theirs with
// ballot ranks. `order` receives a permutation of 0 .. nw-1.
constexpr int kOrderThreads = 1024;
__global__ __launch_bounds__(kOrderThreads) void k_wave_order(const int32_t* __restrict__ cost,
                                                               int64_t nw, int32_t cmax,
                                                               int32_t tail_pct,
                                                               int32_t* __restrict__ order) {
  constexpr int kWaves = kOrderThreads / kWaveLanes;
  __shared__ int32_t hist[kOrderThreads];
  __shared__ int32_t s_short[kWaves], s_long[kWaves];
  const int tid = (int)threadIdx.x;
  const int lane = tid % kWaveLanes, wv = tid / kWaveLanes;
  auto bin = [&](int32_t c) -> int32_t {
    const int64_t b = (int64_t)max(c, 0) * kOrderThreads / ((int64_t)cmax + 1);
    return (int32_t)min(b, (int64_t)kOrderThreads - 1);
  };
  hist[tid] = 0;
  __syncthreads();
  for (int64_t w0 = tid; w0 < nw; w0 += 4 * kOrderThreads) {  // four loads in flight
    int32_t c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t w = w0 + (int64_t)u * kOrderThreads;
      c[u] = w < nw ? cost[w] : -1;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (c[u] >= 0) atomicAdd(&hist[bin(c[u])], 1);
  }
  __syncthreads();
  for (int o = 1; o < kOrderThreads; o <<= 1) {  // inclusive scan of hist
    const int32_t v = tid >= o ? hist[tid - o] : 0;
    __syncthreads();
    hist[tid] += v;
    __syncthreads();
  }
  // bins below thr go last: those whose inclusive count stays within the share
  const int64_t want = nw * tail_pct / 100;
  const int32_t thr = __syncthreads_count((int64_t)hist[tid] <= want);
  const int64_t nlong = nw - (thr > 0 ? hist[thr - 1] : 0);
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (kWaveLanes - lane));
  int64_t short_base = 0, long_base = 0;
  int32_t c_next = tid < nw ? cost[tid] : 0;  // the next round's cost, loaded a round ahead
  for (int64_t base = 0; base < nw; base += kOrderThreads) {
    const int64_t w = base + tid;
    const bool valid = w < nw;
    const int32_t c = c_next;
    if (w + kOrderThreads < nw) c_next = cost[w + kOrderThreads];
    const bool sh = valid && bin(c) < thr;
    const uint64_t ms = __ballot(sh), ml = __ballot(valid && !sh);
    if (lane == 0) {
      s_short[wv] = __popcll(ms);
      s_long[wv] = __popcll(ml);
    }
    __syncthreads();
    int32_t sb = 0, lb = 0, ts = 0, tl = 0;
    for (int k = 0; k < kWaves; ++k) {
      if (k < wv) {
        sb += s_short[k];
        lb += s_long[k];
      }
      ts += s_short[k];
      tl += s_long[k];
    }
    if (sh)
      order[nlong + short_base + sb + __popcll(ms & below)] = (int32_t)w;
    else if (valid)
      order[long_base + lb + __popcll(ml & below)] = (int32_t)w;
    short_base += ts;
    long_base += tl;
    __syncthreads();
  }
}

// skip[q[k]] = v for the n listed queries; an index outside [0, nq) is left
// out and raises the walk-error status bit (the resume kernel clamps it)
__global__ __launch_bounds__(kBlock) void k_mark_queries(const int32_t* __restrict__ q,
                                                          int64_t n, int64_t nq, int32_t v,
                                                          int32_t* __restrict__ skip,
                                                          int32_t* __restrict__ status) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int32_t i = q[k];
  if (i >= 0 && i < nq)
    skip[i] = v;
  else
    atomicOr(status, kStatusStack);
}

// list[from + j] = the query of the j-th saved walk in launch order, j < n
__global__ __launch_bounds__(kBlock) void k_early_list(const SdfSuspended* __restrict__ susp,
                                                        const uint32_t* __restrict__ order,
                                                        int64_t from, int64_t n,
                                                        int32_t* __restrict__ list) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  list[from + j] = susp[order ? (int64_t)order[j] : j].qidx;
}

// ---- mesh set-up -------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_vert_bounds(const double* __restrict__ x,
                                                         const double* __restrict__ y,
                                                         const double* __restrict__ z, int64_t n,
                                                         double* __restrict__ partials) {
  __shared__ double s_red[6][kBlock];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double acc[6] = {kDblMax, kDblMax, kDblMax, -kDblMax, -kDblMax, -kDblMax};
  if (i < n) {
    acc[0] = acc[3] = x[i];
    acc[1] = acc[4] = y[i];
    acc[2] = acc[5] = z[i];
  }
  for (int k = 0; k < 6; ++k) s_red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      for (int k = 0; k < 3; ++k) {
        const double a = s_red[k][threadIdx.x], b = s_red[k][threadIdx.x + off];
        s_red[k][threadIdx.x] = (b < a) ? b : a;
        const double c = s_red[3 + k][threadIdx.x], d = s_red[3 + k][threadIdx.x + off];
        s_red[3 + k][threadIdx.x] = (d > c) ? d : c;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; ++k) partials[6 * (int64_t)blockIdx.x + k] = s_red[k][0];
}

__global__ __launch_bounds__(1024) void k_final_bounds(const double* __restrict__ partials,
                                                        int64_t np, double* __restrict__ out) {
  __shared__ double s_red[6][1024];
  double acc[6] = {kDblMax, kDblMax, kDblMax, -kDblMax, -kDblMax, -kDblMax};
  for (int64_t p = threadIdx.x; p < np; p += 1024)
    for (int k = 0; k < 3; ++k) {
      const double a = partials[6 * p + k];
      if (a < acc[k]) acc[k] = a;
      const double b = partials[6 * p + 3 + k];
      if (b > acc[3 + k]) acc[3 + k] = b;
    }
  for (int k = 0; k < 6; ++k) s_red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      for (int k = 0; k < 3; ++k) {
        const double a = s_red[k][threadIdx.x], b = s_red[k][threadIdx.x + off];
        s_red[k][threadIdx.x] = (b < a) ? b : a;
        const double c = s_red[3 + k][threadIdx.x], d = s_red[3 + k][threadIdx.x + off];
        s_red[3 + k][threadIdx.x] = (d > c) ? d : c;
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; ++k) out[k] = s_red[k][0];
}

// getCellBoundingBox (SignedDistance.hpp:609-633): addPoint of the 3 vertices.
__global__ __launch_bounds__(kBlock) void k_tri_boxes(const double* __restrict__ x,
                                                       const double* __restrict__ y,
                                                       const double* __restrict__ z,
                                                       const int32_t* __restrict__ conn,
                                                       int64_t nt, double* __restrict__ boxes) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nt) return;
  double b[6] = {kDblMax, kDblMax, kDblMax, -kDblMax, -kDblMax, -kDblMax};
  for (int k = 0; k < 3; ++k) {
    const int32_t v = conn[3 * c + k];
    const double p[3] = {x[v], y[v], z[v]};
    for (int d = 0; d < 3; ++d) {
      if (p[d] < b[d]) b[d] = p[d];
      if (p[d] > b[3 + d]) b[3 + d] = p[d];
    }
  }
  for (int k = 0; k < 6; ++k) boxes[6 * c + k] = b[k];
}

// One record per BVH leaf (nleaves >= nt: a single triangle is padded with an
// invalid leaf that traversal never reaches; its record is zero).
__global__ __launch_bounds__(kBlock) void k_gather_tris(const double* __restrict__ x,
                                                         const double* __restrict__ y,
                                                         const double* __restrict__ z,
                                                         const int32_t* __restrict__ conn,
                                                         const int32_t* __restrict__ leaf_ids,
                                                         int64_t nt, int64_t nleaves,
                                                         TriRec* __restrict__ tris) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nleaves) return;
  const int32_t c = leaf_ids[p];
  TriRec r;
  if (c < 0 || c >= nt) {
    for (int k = 0; k < 9; ++k) r.v[k] = 0.0;
    r.id = -1;
    r.pad = 0;
    tris[p] = r;
    return;
  }
  for (int k = 0; k < 3; ++k) {
    const int32_t v = conn[3 * (int64_t)c + k];
    r.v[3 * k] = x[v];
    r.v[3 * k + 1] = y[v];
    r.v[3 * k + 2] = z[v];
  }
  r.id = c;
  r.pad = 0;
  tris[p] = r;
}

// boundaries of sorted keys: start[key] = first index, end[key] = one past
// last. A key >= nseg is not written and raises bit 2 of *status.
// Box of a cell's three or four nodes: addPoint from an invalid box
// (getCellBoundingBox, SignedDistance.hpp:609-633).
__global__ __launch_bounds__(kBlock) void k_cell_boxes(const double* __restrict__ x,
                                                        const double* __restrict__ y,
                                                        const double* __restrict__ z,
                                                        const int32_t* __restrict__ conn,
                                                        const int32_t* __restrict__ off,
                                                        int64_t nc, double* __restrict__ boxes) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double b[6] = {kDblMax, kDblMax, kDblMax, -kDblMax, -kDblMax, -kDblMax};
  for (int32_t k = off[c]; k < off[c + 1]; ++k) {
    const int32_t v = conn[k];
    const double p[3] = {x[v], y[v], z[v]};
    for (int d = 0; d < 3; ++d) {
      if (p[d] < b[d]) b[d] = p[d];
      if (p[d] > b[3 + d]) b[3 + d] = p[d];
    }
  }
  for (int k = 0; k < 6; ++k) boxes[6 * c + k] = b[k];
}

// Two records per BVH leaf: the cell's triangles (n0, n1, n2) and, for a
// quad, (n0, n2, n3), which checkCandidate tests in that order
// (SignedDistance.hpp:649-658). TriRec::id is the cell id; the first
// record's pad is the cell's triangle count (0 for a padding leaf).
__global__ __launch_bounds__(kBlock) void k_gather_cells(const double* __restrict__ x,
                                                          const double* __restrict__ y,
                                                          const double* __restrict__ z,
                                                          const int32_t* __restrict__ conn,
                                                          const int32_t* __restrict__ off,
                                                          const int32_t* __restrict__ leaf_ids,
                                                          int64_t nc, int64_t nleaves,
                                                          TriRec* __restrict__ tris) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nleaves) return;
  const int32_t c = leaf_ids[p];
  const bool real = c >= 0 && c < nc;
  const int32_t b = real ? off[c] : 0;
  const int32_t nn = real ? off[c + 1] - b : 0;
  const int32_t corner[2][3] = {{0, 1, 2}, {0, 2, 3}};
  for (int e = 0; e < 2; ++e) {
    TriRec r;
    const bool have = real && (e == 0 || nn == 4);
    for (int k = 0; k < 3; ++k) {
      const int32_t v = have ? conn[b + corner[e][k]] : 0;
      r.v[3 * k] = have ? x[v] : 0.0;
      r.v[3 * k + 1] = have ? y[v] : 0.0;
      r.v[3 * k + 2] = have ? z[v] : 0.0;
    }
    r.id = real ? c : -1;
    r.pad = (e == 0) ? (real ? (nn == 4 ? 2 : 1) : 0) : 0;
    tris[2 * p + e] = r;
  }
}

__global__ __launch_bounds__(kBlock) void k_segments(const uint32_t* __restrict__ keys, int64_t n,
                                                      int64_t nseg, int32_t* __restrict__ start,
                                                      int32_t* __restrict__ end,
                                                      int32_t* __restrict__ status) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t k = keys[p];
  if ((int64_t)k >= nseg) {
    atomicOr(status, kStatusElem);
    return;
  }
  if (p == 0 || keys[p - 1] != k) start[k] = (int32_t)p;
  if (p == n - 1 || keys[p + 1] != k) end[k] = (int32_t)(p + 1);
}

// Every connectivity entry must name a vertex; bit 4 of *status otherwise.
__global__ __launch_bounds__(kBlock) void k_check_conn(const int32_t* __restrict__ conn, int64_t m,
                                                        int64_t nv, int32_t* __restrict__ status) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m) return;
  const int32_t v = conn[s];
  if (v < 0 || (int64_t)v >= nv) atomicOr(status, 4);
}

__global__ __launch_bounds__(kBlock) void k_iota_conn(const int32_t* __restrict__ conn, int64_t m,
                                                       uint32_t* __restrict__ keys,
                                                       uint32_t* __restrict__ vals) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m) return;
  keys[s] = (uint32_t)conn[s];
  vals[s] = (uint32_t)s;
}

__global__ __launch_bounds__(kBlock) void k_copy_u32(const uint32_t* __restrict__ in, int64_t n,
                                                      int32_t* __restrict__ out) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) out[s] = (int32_t)in[s];
}

// ---- adjoint -----------------------------------------------------------------
// Sort keys of the adjoint: the winning triangle, or nt for a query without
// one (elem -1), which sorts after every triangle; any other id outside
// [0, nt) sets bit kStatusElem of *status and also takes key nt.
__global__ __launch_bounds__(kBlock) void k_elem_keys(const int32_t* __restrict__ elem, int64_t nq,
                                                       int64_t nt, uint32_t* __restrict__ keys,
                                                       uint32_t* __restrict__ vals,
                                                       int32_t* __restrict__ status) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const int32_t e = elem[i];
  const bool in_range = e >= 0 && (int64_t)e < nt;
  if (!in_range && e != -1) atomicOr(status, kStatusElem);
  keys[i] = in_range ? (uint32_t)e : (uint32_t)nt;
  vals[i] = (uint32_t)i;
}

// One lane per query, in element order (position p of the sorted list): the
// query gradient, and the closest point's barycentric weights w and the
// scaled direction d that the per-triangle sums read (terms[6p .. 6p + 5]).
__global__ __launch_bounds__(kBlock) void k_bwd_query(
    const double* __restrict__ vx, const double* __restrict__ vy, const double* __restrict__ vz,
    const int32_t* __restrict__ conn, int64_t nt, const uint32_t* __restrict__ keys,
    const uint32_t* __restrict__ qidx, int64_t nq, const double* __restrict__ qx,
    const double* __restrict__ qy, const double* __restrict__ qz, int64_t qs,
    const double* __restrict__ dist, const float* __restrict__ gout, double* __restrict__ grad_q,
    double* __restrict__ terms) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nq) return;
  const int64_t c = (int64_t)keys[p];
  if (c >= nt) {  // no closest triangle: the distance is constant in the query
    if (grad_q) {
      const int64_t i = (int64_t)qidx[p];
      grad_q[3 * i] = 0.0;
      grad_q[3 * i + 1] = 0.0;
      grad_q[3 * i + 2] = 0.0;
    }
    return;
  }
  double T[9];
  for (int k = 0; k < 3; ++k) {
    const int32_t v = conn[3 * c + k];
    T[3 * k] = vx[v];
    T[3 * k + 1] = vy[v];
    T[3 * k + 2] = vz[v];
  }
  const int64_t i = (int64_t)qidx[p];
  const double P[3] = {qx[i * qs], qy[i * qs], qz[i * qs]};
  double cp[3], w[3];
  closest_point<true>(P, T, 1e-12, cp, w);
  const double sgn = signbit(dist[i]) ? -1.0 : 1.0;
  const double r0 = P[0] - cp[0], r1 = P[1] - cp[1], r2 = P[2] - cp[2];
  const double len = sqrt(dot3(r0, r1, r2, r0, r1, r2));
  double u0, u1, u2;
  if (len > 0.0) {
    const double inv = 1.0 / len;
    u0 = r0 * inv;
    u1 = r1 * inv;
    u2 = r2 * inv;
  } else {  // on the surface: the face normal
    double nrm[3], un[3];
    tri_normal(T, nrm);
    unit3(nrm, un);
    u0 = un[0];
    u1 = un[1];
    u2 = un[2];
  }
  const double g = (double)gout[i] * sgn;
  const double d0 = g * u0, d1 = g * u1, d2 = g * u2;
  if (grad_q) {
    grad_q[3 * i] = d0;
    grad_q[3 * i + 1] = d1;
    grad_q[3 * i + 2] = d2;
  }
  double* t = terms + 6 * p;
  t[0] = w[0];
  t[1] = w[1];
  t[2] = w[2];
  t[3] = d0;
  t[4] = d1;
  t[5] = d2;
}

// One wave per triangle: the per-corner partial sums (9 doubles per
// triangle) over its queries in ascending index order, from k_bwd_query's
// terms. The lanes load 64 queries' terms at a time and form their products;
// lane k < 9 then adds component k in query order. The sums are the
// one-lane sequential sums, bit for bit, while a triangle reached by many
// queries (hundreds, where queries share a vertex's normal cone) costs a few
// coalesced loads instead of a chain of hundreds.
constexpr int kBwdWave = 64;
__global__ __launch_bounds__(kBwdWave) void k_bwd_tri(int64_t nt,
                                                       const int32_t* __restrict__ seg_start,
                                                       const int32_t* __restrict__ seg_end,
                                                       const double* __restrict__ terms,
                                                       double* __restrict__ part) {
  __shared__ double s_pr[9][kBwdWave];
  const int64_t c = blockIdx.x;
  const int lane = (int)threadIdx.x;
  if (c >= nt) return;
  const int32_t b = seg_start[c], e = seg_end[c];
  double acc = 0.0;  // lane k < 9: component k (corner k / 3, axis k % 3)
  for (int32_t base = b; base < e; base += kBwdWave) {
    const int32_t p = base + lane;
    if (p < e) {
      const double* t = terms + 6 * (int64_t)p;
      const double w[3] = {t[0], t[1], t[2]};
      const double d0 = t[3], d1 = t[4], d2 = t[5];
      for (int k = 0; k < 3; ++k) {
        s_pr[3 * k][lane] = -(w[k] * d0);
        s_pr[3 * k + 1][lane] = -(w[k] * d1);
        s_pr[3 * k + 2][lane] = -(w[k] * d2);
      }
    }
    __syncthreads();
    const int32_t m = (e - base < kBwdWave) ? e - base : kBwdWave;
    if (lane < 9)
      for (int32_t l = 0; l < m; ++l) acc += s_pr[lane][l];
    __syncthreads();
  }
  if (lane < 9) part[9 * c + lane] = acc;
}

// One lane per vertex: sum of its incident corner partials, ascending triangle.
__global__ __launch_bounds__(kBlock) void k_bwd_vert(const int32_t* __restrict__ vstart,
                                                      const int32_t* __restrict__ vend,
                                                      const int32_t* __restrict__ slots,
                                                      int64_t nv, const double* __restrict__ part,
                                                      double* __restrict__ grad_v) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  const int32_t b = vstart[v], e = vend[v];
  for (int32_t k = b; k < e; ++k) {
    const int64_t s = slots[k];  // 3*tri + corner
    const double* p = part + 3 * s;
    a0 += p[0];
    a1 += p[1];
    a2 += p[2];
  }
  grad_v[3 * v] = a0;
  grad_v[3 * v + 1] = a1;
  grad_v[3 * v + 2] = a2;
}

template <class T>
int dmalloc(T** p, size_t count) {
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (count ? count : 1));
  if (e != hipSuccess) {
    set_error("hipMalloc of %zu bytes failed: %s", sizeof(T) * count, hipGetErrorString(e));
    return AXSQ_ERR_NOMEM;
  }
  return AXSQ_OK;
}

}  // namespace

// Rejects connectivity that names a vertex outside [0, nverts).
int sdf_check_conn(const axsq_sdf* h, hipStream_t s) {
  int32_t* flag = nullptr;
  AXSQ_TRY(dmalloc(&flag, 1));
  int rc = fill_i32(flag, 0, 1, s);
  if (rc == AXSQ_OK) {
    k_check_conn<<<grid_for(h->nconn), kBlock, 0, s>>>(h->conn, h->nconn, h->nverts, flag);
    int32_t hf = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&hf, flag, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      set_error("connectivity check failed to run");
      rc = AXSQ_ERR_HIP;
    } else if (hf) {
      set_error("cell connectivity names a vertex outside [0, %lld)", (long long)h->nverts);
      rc = AXSQ_ERR_INVALID_ARG;
    }
  }
  (void)hipFree(flag);
  return rc;
}

// Mesh bounds, triangle boxes, BVH, leaf-ordered triangles.
int sdf_setup(axsq_sdf* h, hipStream_t s) {
  Workspace& ws = workspace();
  const int64_t nv = h->nverts, nt = h->ntris;
  const int64_t np = grid_for(nv);
  AXSQ_TRY(ws.tiles.ensure(sizeof(double) * 6 * (size_t)(np + 1)));
  double* part = ws.tiles.as<double>();
  k_vert_bounds<<<(unsigned)np, kBlock, 0, s>>>(h->vx, h->vy, h->vz, nv, part);
  AXSQ_CHECK_HIP(hipGetLastError());
  k_final_bounds<<<1, 1024, 0, s>>>(part, np, part + 6 * np);
  AXSQ_CHECK_HIP(hipGetLastError());
  AXSQ_CHECK_HIP(hipMemcpyAsync(h->domain, part + 6 * np, 6 * sizeof(double),
                                hipMemcpyDeviceToHost, s));
  AXSQ_CHECK_HIP(hipStreamSynchronize(s));
  // triangle boxes live in a temporary allocation: the BVH build uses the
  // whole workspace for its own scratch
  double* boxes = nullptr;
  AXSQ_TRY(dmalloc(&boxes, 6 * (size_t)nt));
  if (h->cells)
    k_cell_boxes<<<grid_for(nt), kBlock, 0, s>>>(h->vx, h->vy, h->vz, h->conn, h->cell_off, nt,
                                                 boxes);
  else
    k_tri_boxes<<<grid_for(nt), kBlock, 0, s>>>(h->vx, h->vy, h->vz, h->conn, nt, boxes);
  hipError_t e = hipGetLastError();
  int rc = (e == hipSuccess) ? AXSQ_OK : AXSQ_ERR_HIP;
  if (rc == AXSQ_OK) rc = bvh_build_into(h->bvh, boxes, nt, 1.000123, h->fp_mode, s);
  hipFree(boxes);
  if (e != hipSuccess) set_error("k_tri_boxes: %s", hipGetErrorString(e));
  AXSQ_TRY(rc);
  const int64_t nl = h->bvh->n;
  if (!h->tris) AXSQ_TRY(dmalloc(&h->tris, (size_t)(h->cells ? 2 * nl : nl)));
  if (h->cells)
    k_gather_cells<<<grid_for(nl), kBlock, 0, s>>>(h->vx, h->vy, h->vz, h->conn, h->cell_off,
                                                   h->bvh->leaf_ids, nt, nl, h->tris);
  else
    k_gather_tris<<<grid_for(nl), kBlock, 0, s>>>(h->vx, h->vy, h->vz, h->conn, h->bvh->leaf_ids,
                                                  nt, nl, h->tris);
  AXSQ_CHECK_HIP(hipGetLastError());
  AXSQ_CHECK_HIP(hipStreamSynchronize(s));
  return AXSQ_OK;
}

// vertex -> incident (3*tri + corner) slots, ascending, by a stable sort of
// the connectivity.
int sdf_build_v2t(axsq_sdf* h, hipStream_t s) {
  Workspace& ws = workspace();
  const int64_t m = 3 * h->ntris;
  AXSQ_TRY(ws.k0.ensure(sizeof(uint32_t) * m));
  AXSQ_TRY(ws.v0.ensure(sizeof(uint32_t) * m));
  AXSQ_TRY(ws.k1.ensure(sizeof(uint32_t) * m));
  AXSQ_TRY(ws.v1.ensure(sizeof(uint32_t) * m));
  if (!h->v2t_slot) {
    AXSQ_TRY(dmalloc(&h->v2t_slot, (size_t)m));
    AXSQ_TRY(dmalloc(&h->v2t_start, (size_t)h->nverts));
    AXSQ_TRY(dmalloc(&h->v2t_end, (size_t)h->nverts));
  }
  k_iota_conn<<<grid_for(m), kBlock, 0, s>>>(h->conn, m, ws.k0.as<uint32_t>(),
                                             ws.v0.as<uint32_t>());
  AXSQ_CHECK_HIP(hipGetLastError());
  uint32_t *sk, *sv;
  AXSQ_TRY(radix_sort_pairs(ws.k0.as<uint32_t>(), ws.v0.as<uint32_t>(), ws.k1.as<uint32_t>(),
                            ws.v1.as<uint32_t>(), m, bits_for((uint64_t)h->nverts), ws, s, &sk,
                            &sv));
  AXSQ_TRY(fill_i32(h->v2t_start, 0, h->nverts, s));
  AXSQ_TRY(fill_i32(h->v2t_end, 0, h->nverts, s));
  k_segments<<<grid_for(m), kBlock, 0, s>>>(sk, m, h->nverts, h->v2t_start, h->v2t_end,
                                          h->bvh->status);
  AXSQ_CHECK_HIP(hipGetLastError());
  k_copy_u32<<<grid_for(m), kBlock, 0, s>>>(sv, m, h->v2t_slot);
  AXSQ_CHECK_HIP(hipGetLastError());
  AXSQ_CHECK_HIP(hipStreamSynchronize(s));
  return AXSQ_OK;
}

// A side stream and fork/join events for the early walks, one set per host
// thread and device (default priority, non-blocking).
struct SideStream {
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};
static SideStream* side_stream() {
  constexpr int kMaxDev = 64;
  thread_local SideStream sides[kMaxDev];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) return nullptr;
  SideStream& ss = sides[dev];
  if (!ss.stream) {
    if (hipStreamCreateWithFlags(&ss.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ss.join, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      ss = SideStream{};
      return nullptr;
    }
  }
  return &ss;
}

int sdf_eval_launch(const axsq_sdf* h, const double* x, const double* y, const double* z,
                    int64_t stride, int64_t nq, int fp_mode, double* dist, double* cp,
                    double* normal, int32_t* elem, bool stats, hipStream_t s) {
  Workspace& ws = workspace();
  const axsq_bvh* b = h->bvh;
  SdfArgs a{};
  a.nodes = b->nodes;
  a.nodes32 = b->nodes32;
  a.ranges32 = b->ranges32;
  a.tris = h->tris;
  a.qx = x;
  a.qy = y;
  a.qz = z;
  a.qs = stride;
  a.nq = nq;
  a.ninner = (int32_t)(b->n - 1);
  a.order = nullptr;
  a.susp_order = nullptr;
  for (int k = 0; k < 6; ++k) a.dom[k] = h->domain[k];
  a.watertight = h->watertight;
  a.compute_sign = h->compute_sign;
  a.dist = dist;
  a.cp = cp;
  a.normal = normal;
  a.elem = elem;
  a.status = b->status;
  a.stats = h->stats_dev;
  if (h->sort_queries && nq > 4096) {
    AXSQ_TRY(ws.k0.ensure(sizeof(uint32_t) * nq));
    AXSQ_TRY(ws.v0.ensure(sizeof(uint32_t) * nq));
    AXSQ_TRY(ws.k1.ensure(sizeof(uint32_t) * nq));
    AXSQ_TRY(ws.v1.ensure(sizeof(uint32_t) * nq));
    double inv[3];
    for (int d = 0; d < 3; ++d) {
      const double ext = h->domain[3 + d] - h->domain[d];
      inv[d] = ext > 0.0 ? 1.0 / ext : 0.0;
    }
    ProfRange pq("sdf_query_order", s);
    k_query_codes<<<grid_for(nq), kBlock, 0, s>>>(x, y, z, stride, nq, h->domain[0],
                                                  h->domain[1], h->domain[2], inv[0], inv[1],
                                                  inv[2], ws.k0.as<uint32_t>(),
                                                  ws.v0.as<uint32_t>());
    AXSQ_CHECK_HIP(hipGetLastError());
    uint32_t *sk, *sv;
    AXSQ_TRY(radix_sort_pairs(ws.k0.as<uint32_t>(), ws.v0.as<uint32_t>(), ws.k1.as<uint32_t>(),
                              ws.v1.as<uint32_t>(), nq, 30, ws, s, &sk, &sv));
    pq.stop();
    a.order = sv;
  }
  const bool cells = h->cells;  // quad cells: FP64 nodes, two records per leaf
  const bool mixed = (fp_mode == AXSQ_FP32_TRAVERSAL) && !cells;
  // room for suspended walks (360 B each): as many as the previous call on
  // this handle suspended, plus a quarter and 4,096; on the first call an
  // eighth of the queries (at least 65,536); never more than the queries or
  // 2^24 (6.0 GB). A walk with no room left finishes in the first pass, and
  // the next call makes room for it.
  const int64_t full = std::min<int64_t>(nq, int64_t(1) << 24);
  int64_t cap;
  if (h->susp_cap_fixed >= 0)
    cap = std::min<int64_t>(full, h->susp_cap_fixed);
  else if (h->susp_demand < 0)
    cap = std::min<int64_t>(full, std::max<int64_t>(65536, nq / 8));
  else
    cap = std::min<int64_t>(full, h->susp_demand + h->susp_demand / 4 + 4096);
  cap = std::max<int64_t>(cap, 1);
  AXSQ_TRY(ws.susp.ensure(sizeof(SdfSuspended) * (size_t)cap));
  AXSQ_TRY(ws.redo.ensure(sizeof(int32_t) * (size_t)nq));
  AXSQ_TRY(ws.cnt.ensure(sizeof(uint32_t) * 2));
  a.susp = ws.susp.as<SdfSuspended>();
  a.nsusp = ws.cnt.as<uint32_t>();
  a.redo = ws.redo.as<int32_t>();
  a.nredo = ws.cnt.as<uint32_t>() + 1;
  a.cap = cap;
  a.budget = h->budget;
  AXSQ_TRY(fill_i32(reinterpret_cast<int32_t*>(a.nsusp), 0, 2, s));
  ProfRange pr(mixed ? "sdf_eval_fp32" : "sdf_eval_fp64", s);
  // Early walks: the previous call on the same query arrays listed its
  // longest walks (the first of its second pass's launch order). Their
  // queries are walked here from the start, one wave each as in the second
  // pass, on a side stream beside the first pass, which skips them; so the
  // longest walks no longer wait for the first pass to end. Which queries
  // these are changes speed only: each query is still walked once, in full.
  axsq_sdf* hm = const_cast<axsq_sdf*>(h);
  const uint64_t qkey[5] = {(uint64_t)(uintptr_t)x, (uint64_t)(uintptr_t)y,
                            (uint64_t)(uintptr_t)z, (uint64_t)stride, (uint64_t)nq};
  const int64_t kEarly = cells ? 0 : h->early_walks;
  if (kEarly == 0 || memcmp(qkey, hm->early_key, sizeof(qkey)) != 0) {
    hm->nearly = 0;
    memcpy(hm->early_key, qkey, sizeof(qkey));
  }
  const int64_t nearly = std::min<int64_t>(hm->nearly, kEarly);
  SideStream* side = nearly > 0 ? side_stream() : nullptr;
  // launch order of the waves from the previous call's steps (in query order)
  const bool wave_on = h->wave_tail > 0 && a.order != nullptr;
  if ((kEarly > 0 || wave_on) && !hm->early_done)
    AXSQ_CHECK_HIP(hipEventCreateWithFlags(&hm->early_done, hipEventDisableTiming));
  // On every way out of this call once the early walks are forked (an error
  // too): clear the marks if pass 1 has not, make the caller's stream wait
  // for the side stream, and record where this call last touches the list
  // and marks.
  struct EarlyExit {
    hipStream_t s = nullptr;
    SideStream* side = nullptr;
    axsq_sdf* hm = nullptr;
    int64_t n = 0, nq = 0;
    int32_t* status = nullptr;
    bool marked = false, record = false;
    ~EarlyExit() {
      if (marked)
        k_mark_queries<<<grid_for(n), kBlock, 0, s>>>(hm->early_q, n, nq, 0, hm->skip, status);
      if (side) (void)hipStreamWaitEvent(s, side->join, 0);
      if (record && hm->early_done) (void)hipEventRecord(hm->early_done, s);
    }
  } early_exit;
  early_exit.s = s;
  early_exit.hm = hm;
  early_exit.n = nearly;
  early_exit.nq = nq;
  early_exit.status = b->status;
  early_exit.record = kEarly > 0 || wave_on;
  const bool mixed2_early = mixed && h->mixed_pass2_fp32;
  if (side) {
    // the previous call may have run on another stream: its last writes of
    // the list and marks come first
    if (hm->early_done) AXSQ_CHECK_HIP(hipStreamWaitEvent(s, hm->early_done, 0));
    if (hm->skip_cap < nq) {
      (void)hipFree(hm->skip);
      hm->skip = nullptr;
      hm->skip_cap = 0;
      AXSQ_TRY(dmalloc(&hm->skip, (size_t)nq));
      hm->skip_cap = nq;
      AXSQ_TRY(fill_i32(hm->skip, 0, nq, s));
    }
    k_mark_queries<<<grid_for(nearly), kBlock, 0, s>>>(hm->early_q, nearly, nq, 1, hm->skip,
                                                            b->status);
    early_exit.marked = true;
    AXSQ_CHECK_HIP(hipGetLastError());
    AXSQ_CHECK_HIP(hipEventRecord(side->fork, s));
    AXSQ_CHECK_HIP(hipStreamWaitEvent(side->stream, side->fork, 0));
    SdfArgs ae = a;
    ae.redo = hm->early_q;  // walks from the start, like those of overflowed term lists
    ae.susp = nullptr;
    ae.susp_order = nullptr;
    ae.nhot = 0;
    ae.hot_leaf = ae.hot_node = nullptr;
    const unsigned ge = (unsigned)nearly;
    if (mixed2_early)
      k_sdf_resume_wave<true, false><<<ge, kWaveLanes, 0, side->stream>>>(ae, 0, nearly);
    else if (stats)
      k_sdf_resume_wave<false, true><<<ge, kWaveLanes, 0, side->stream>>>(ae, 0, nearly);
    else
      k_sdf_resume_wave<false, false><<<ge, kWaveLanes, 0, side->stream>>>(ae, 0, nearly);
    AXSQ_CHECK_HIP(hipGetLastError());
    AXSQ_CHECK_HIP(hipEventRecord(side->join, side->stream));
    early_exit.side = side;
    a.skip = hm->skip;
  }
  if (wave_on) {
    // the previous call may have run on another stream: its first pass
    // wrote the steps read here and rewritten by this call's
    if (hm->early_done) AXSQ_CHECK_HIP(hipStreamWaitEvent(s, hm->early_done, 0));
    const bool same = hm->wsteps_valid && memcmp(qkey, hm->wave_key, sizeof(qkey)) == 0;
    if (hm->wsteps_cap < nq) {
      (void)hipFree(hm->wsteps);
      hm->wsteps = nullptr;
      hm->wsteps_cap = 0;
      hm->wsteps_valid = false;
      AXSQ_TRY(dmalloc(&hm->wsteps, (size_t)nq));
      hm->wsteps_cap = nq;
    }
    if (same) {
      const int64_t nw = (nq + kWaveLanes - 1) / kWaveLanes;
      AXSQ_TRY(ws.wave_cost.ensure(sizeof(int32_t) * (size_t)nw));
      AXSQ_TRY(ws.wave_order.ensure(sizeof(int32_t) * (size_t)nw));
      k_wave_cost<<<grid_for(nw * kWaveLanes), kBlock, 0, s>>>(hm->wsteps, nq, nw,
                                                               ws.wave_cost.as<int32_t>());
      AXSQ_CHECK_HIP(hipGetLastError());
      k_wave_order<<<1, kOrderThreads, 0, s>>>(ws.wave_cost.as<int32_t>(), nw, h->budget,
                                               h->wave_tail, ws.wave_order.as<int32_t>());
      AXSQ_CHECK_HIP(hipGetLastError());
      a.wave_order = ws.wave_order.as<int32_t>();
    }
    a.wsteps = hm->wsteps;
    memcpy(hm->wave_key, qkey, sizeof(qkey));
    hm->wsteps_valid = true;
  }
  ProfRange p1(mixed ? "sdf_eval_fp32_pass1" : "sdf_eval_fp64_pass1", s);
  const unsigned g = grid_for(nq, kSdfBlock);
  if (cells) {
    if (stats)
      k_sdf_eval<false, true, true><<<g, kSdfBlock, 0, s>>>(a);
    else
      k_sdf_eval<false, false, true><<<g, kSdfBlock, 0, s>>>(a);
  } else if (mixed) {
    k_sdf_eval<true, false><<<g, kSdfBlock, 0, s>>>(a);
  } else if (stats) {
    k_sdf_eval<false, true><<<g, kSdfBlock, 0, s>>>(a);
  } else {
    k_sdf_eval<false, false><<<g, kSdfBlock, 0, s>>>(a);
  }
  AXSQ_CHECK_HIP(hipGetLastError());
  p1.stop();
  if (early_exit.marked) {  // clear the marks for the next call
    early_exit.marked = false;
    k_mark_queries<<<grid_for(nearly), kBlock, 0, s>>>(hm->early_q, nearly, nq, 0, hm->skip,
                                                            b->status);
    AXSQ_CHECK_HIP(hipGetLastError());
  }
  uint32_t counts[2] = {0, 0};
  AXSQ_CHECK_HIP(hipMemcpyAsync(counts, a.nsusp, sizeof(counts), hipMemcpyDeviceToHost, s));
  AXSQ_CHECK_HIP(hipStreamSynchronize(s));
  const int64_t nres = std::min<int64_t>(counts[0], cap);
  const int64_t nredo = std::min<int64_t>(counts[1], nq);
  const_cast<axsq_sdf*>(h)->last_suspended = nres + nredo;
  hm->last_early = side ? nearly : 0;
  const_cast<axsq_sdf*>(h)->susp_demand = counts[0];  // saved or not
  ProfRange p2(mixed ? "sdf_eval_fp32_pass2" : "sdf_eval_fp64_pass2", s);
  a.susp_order = nullptr;
  if (nres > 64) {  // the query order is no longer needed: reuse its buffers
    AXSQ_TRY(ws.k0.ensure(sizeof(uint32_t) * nres));
    AXSQ_TRY(ws.v0.ensure(sizeof(uint32_t) * nres));
    AXSQ_TRY(ws.k1.ensure(sizeof(uint32_t) * nres));
    AXSQ_TRY(ws.v1.ensure(sizeof(uint32_t) * nres));
    k_susp_keys<<<grid_for(nres), kBlock, 0, s>>>(a.susp, nres, ws.k0.as<uint32_t>(),
                                                  ws.v0.as<uint32_t>());
    AXSQ_CHECK_HIP(hipGetLastError());
    uint32_t *sk, *sv;
    AXSQ_TRY(radix_sort_pairs(ws.k0.as<uint32_t>(), ws.v0.as<uint32_t>(), ws.k1.as<uint32_t>(),
                              ws.v1.as<uint32_t>(), nres, 32, ws, s, &sk, &sv));
    a.susp_order = sv;
  }
  // the longest walks of this call (the first of the launch order) join the
  // list for the next call on these queries
  if (kEarly > 0) {
    const int64_t from = side ? nearly : 0;
    const int64_t nadd = std::min<int64_t>(nres, kEarly - from);
    if (nadd > 0) {
      if (!hm->early_q) AXSQ_TRY(dmalloc(&hm->early_q, (size_t)kEarly));
      k_early_list<<<grid_for(nadd), kBlock, 0, s>>>(a.susp, a.susp_order, from, nadd,
                                                      hm->early_q);
      AXSQ_CHECK_HIP(hipGetLastError());
    }
    hm->nearly = from + std::max<int64_t>(nadd, 0);
  }
  // the first walks of the launch order (the farthest from the surface, which
  // run longest) read tables of their leaf records and node keys
  a.nhot = 0;
  a.hot_leaf = a.hot_node = nullptr;
  // The FP32-traversal form's second pass resumes the saved walks on the FP64
  // nodes unless the handle asks for the FP32 ones: pruning from a saved
  // state is exact either way, and on gfx950 the FP64 resume kernel runs 3
  // waves per SIMD where the FP32 one, with its two box tests, fits 2.
  const bool mixed2 = mixed && h->mixed_pass2_fp32;
  const int64_t nleaf = (int64_t)a.ninner + 1;
  const size_t hot_bytes = sizeof(HotLeaf) * nleaf +
                           (mixed2 ? sizeof(HotNode<float>) : sizeof(HotNode<double>)) * a.ninner;
  const int64_t nhot = std::min<int64_t>({nres, (int64_t)AXSQ_HOT_WALKS,
                                          (int64_t)((size_t(2) << 30) / hot_bytes)});
  if (nhot > 0 && !cells) {
    AXSQ_TRY(ws.hot.ensure(hot_bytes * (size_t)nhot));
    a.nhot = nhot;
    a.hot_leaf = ws.hot.as<char>();
    a.hot_node = ws.hot.as<char>() + sizeof(HotLeaf) * nleaf * (size_t)nhot;
    const dim3 gh(grid_for(nleaf + a.ninner), (unsigned)nhot);
    if (mixed2)
      k_hot_tables<true><<<gh, kBlock, 0, s>>>(a);
    else
      k_hot_tables<false><<<gh, kBlock, 0, s>>>(a);
    AXSQ_CHECK_HIP(hipGetLastError());
  }
  if (nres + nredo > 0) {
    const unsigned g2 = (unsigned)(nres + nredo);
    if (cells) {
      if (stats)
        k_sdf_resume_wave<false, true, true><<<g2, kWaveLanes, 0, s>>>(a, nres, nredo);
      else
        k_sdf_resume_wave<false, false, true><<<g2, kWaveLanes, 0, s>>>(a, nres, nredo);
    } else if (mixed2) {
      k_sdf_resume_wave<true, false><<<g2, kWaveLanes, 0, s>>>(a, nres, nredo);
    } else if (stats) {
      k_sdf_resume_wave<false, true><<<g2, kWaveLanes, 0, s>>>(a, nres, nredo);
    } else {
      k_sdf_resume_wave<false, false><<<g2, kWaveLanes, 0, s>>>(a, nres, nredo);
    }
    AXSQ_CHECK_HIP(hipGetLastError());
  }
  if (early_exit.side) {
    early_exit.side = nullptr;
    AXSQ_CHECK_HIP(hipStreamWaitEvent(s, side->join, 0));
  }
  if (early_exit.record) {  // the list and marks are settled once the stream gets here
    early_exit.record = false;
    AXSQ_CHECK_HIP(hipEventRecord(hm->early_done, s));
  }
  p2.stop();
  pr.stop();
  return AXSQ_OK;
}

int sdf_backward_launch(const axsq_sdf* h, const double* x, const double* y, const double* z,
                        int64_t stride, int64_t nq, const double* dist, const int32_t* elem,
                        const float* gout, double* grad_q, double* grad_v, hipStream_t s) {
  Workspace& ws = workspace();
  const int64_t nt = h->ntris;
  AXSQ_TRY(ws.k0.ensure(sizeof(uint32_t) * nq));
  AXSQ_TRY(ws.v0.ensure(sizeof(uint32_t) * nq));
  AXSQ_TRY(ws.k1.ensure(sizeof(uint32_t) * nq));
  AXSQ_TRY(ws.v1.ensure(sizeof(uint32_t) * nq));
  // segments of the nt triangles and of key nt (queries without a triangle;
  // 2*(nt+1) int32), partials (9*nt doubles), per-query terms (6*nq doubles)
  const size_t seg_bytes = (sizeof(int32_t) * 2 * (nt + 1) + 255) & ~size_t(255);
  const size_t part_bytes = sizeof(double) * 9 * nt;
  AXSQ_TRY(ws.flag.ensure(seg_bytes + part_bytes + sizeof(double) * 6 * nq));
  int32_t* seg_start = ws.flag.as<int32_t>();
  int32_t* seg_end = seg_start + nt + 1;
  double* part = reinterpret_cast<double*>(ws.flag.as<char>() + seg_bytes);
  double* terms = reinterpret_cast<double*>(ws.flag.as<char>() + seg_bytes + part_bytes);
  ProfRange pr("sdf_backward", s);
  k_elem_keys<<<grid_for(nq), kBlock, 0, s>>>(elem, nq, nt, ws.k0.as<uint32_t>(),
                                              ws.v0.as<uint32_t>(), h->bvh->status);
  AXSQ_CHECK_HIP(hipGetLastError());
  uint32_t *sk, *sv;
  AXSQ_TRY(radix_sort_pairs(ws.k0.as<uint32_t>(), ws.v0.as<uint32_t>(), ws.k1.as<uint32_t>(),
                            ws.v1.as<uint32_t>(), nq, bits_for((uint64_t)nt), ws, s, &sk, &sv));
  AXSQ_TRY(fill_i32(seg_start, 0, 2 * (nt + 1), s));
  k_segments<<<grid_for(nq), kBlock, 0, s>>>(sk, nq, nt + 1, seg_start, seg_end, h->bvh->status);
  AXSQ_CHECK_HIP(hipGetLastError());
  k_bwd_query<<<grid_for(nq), kBlock, 0, s>>>(h->vx, h->vy, h->vz, h->conn, nt, sk, sv, nq, x, y,
                                              z, stride, dist, gout, grad_q, terms);
  AXSQ_CHECK_HIP(hipGetLastError());
  k_bwd_tri<<<(unsigned)nt, kBwdWave, 0, s>>>(nt, seg_start, seg_end, terms, part);
  AXSQ_CHECK_HIP(hipGetLastError());
  if (grad_v) {
    k_bwd_vert<<<grid_for(h->nverts), kBlock, 0, s>>>(h->v2t_start, h->v2t_end, h->v2t_slot,
                                                      h->nverts, part, grad_v);
    AXSQ_CHECK_HIP(hipGetLastError());
  }
  pr.stop();
  return AXSQ_OK;
}

}  // namespace axsq
