// axsq_mesh.hip -- triangle-triangle intersection and the BVH mesh
// self-intersection test (quest::findTriMeshIntersectionsBVH) for gfx950.
//
// The reference (quest/detail/MeshTester_detail.hpp:159-341) builds corner
// boxes and degenerate flags, a BVH over the boxes, runs findBoundingBoxes
// for every box (count, scan, fill of all candidates), keeps the pairs i < j
// and tests them. Here the triangle test is the traversal's leaf action
// (the custom leaf action of quest/examples/quest_bvh_two_pass.cpp:144-232):
// one lane per triangle walks the tree with its box and tests each reached
// triangle j > i on the spot, keeping the first kPairSlots hits. A scan of the
// hit counts places every triangle's pairs; triangles with more hits walk
// again to write them all. The candidate lists are never stored. The output
// order is the sequential path's: pairs grouped by the first triangle, the
// second in the first's BVH traversal order; degenerate indices ascending.
#include <hip/hip_runtime.h>

#include "axsq_internal.hpp"
#include "axsq_traverse.hpp"
#include "axsq_tritri.hpp"

namespace axsq {
namespace {

__device__ __forceinline__ V3 load_v3(const double* p) { return {p[0], p[1], p[2]}; }

// Triangle corners (9 doubles, A B C), the corner box (BoundingBox{A, B, C},
// compute_bounding_box.hpp:110-114, BoundingBox.hpp:405-422,464-482) and the
// degenerate flag (Triangle::degenerate, Triangle.hpp:104-109,327-330).
__global__ __launch_bounds__(kBlock) void k_tri_setup(const double* __restrict__ x,
                                                      const double* __restrict__ y,
                                                      const double* __restrict__ z,
                                                      const int32_t* __restrict__ conn,
                                                      int64_t n, int64_t nverts,
                                                      double* __restrict__ tri,
                                                      double* __restrict__ boxes,
                                                      int32_t* __restrict__ degen,
                                                      int32_t* __restrict__ status) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double t[9];
  for (int k = 0; k < 3; ++k) {
    int32_t v = conn[3 * c + k];
    if (v < 0 || v >= nverts) {
      atomicOr(status, 2);
      v = 0;
    }
    t[3 * k] = x[v];
    t[3 * k + 1] = y[v];
    t[3 * k + 2] = z[v];
  }
  double b[6] = {t[0], t[1], t[2], t[0], t[1], t[2]};
  for (int corner = 1; corner < 3; ++corner)
    for (int d = 0; d < 3; ++d) {
      const double p = t[3 * corner + d];
      if (p < b[d]) b[d] = p;
      if (p > b[3 + d]) b[3 + d] = p;
    }
  double nrm[3];
  tri_normal(t, nrm);
  degen[c] = near_eq(0.5 * sqrt(dot3(nrm[0], nrm[1], nrm[2], nrm[0], nrm[1], nrm[2])), 0.0, 1.0e-12)
                 ? 1
                 : 0;
  for (int k = 0; k < 9; ++k) tri[9 * c + k] = t[k];
  for (int k = 0; k < 6; ++k) boxes[6 * c + k] = b[k];
}

constexpr int kPairSlots = 4;  // hits kept per triangle by the first walk

struct IsectArgs {
  const Node64* nodes;
  int32_t ninner;
  const int32_t* leaf_ids;
  const double* tri;    // 9 doubles per triangle
  const double* boxes;  // 6 doubles per triangle (the unscaled corner box)
  int64_t n;
  double eps;
  int32_t* npair;  // hits per triangle
  int32_t* slots;  // kPairSlots per triangle
  const int32_t* poff;
  int32_t* first;
  int32_t* second;
  int32_t* status;
};

// One walk of triangle i's box through the tree (BoundingBox::intersectsWith,
// BVH.hpp:558-560, as findBoundingBoxes tests it) that calls hit(j) for each
// reached triangle j > i that intersects triangle i (boundary excluded).
template <class Hit>
__device__ __forceinline__ void isect_walk(const IsectArgs& a, int64_t i, lds_i32* stk, Hit&& hit) {
  double prim[6];
  for (int k = 0; k < 6; ++k) prim[k] = a.boxes[6 * i + k];
  const double* ti = a.tri + 9 * i;
  const V3 a0 = load_v3(ti), a1 = load_v3(ti + 3), a2 = load_v3(ti + 6);
  auto no_pref = [](const auto*, const auto*) { return false; };
  const int32_t v = traverse_voted<Node64, double>(
      a.nodes, a.ninner, stk,
      [&](const double* b) {
        return (prim[3] >= b[0] && prim[0] <= b[3]) && (prim[4] >= b[1] && prim[1] <= b[4]) &&
               (prim[5] >= b[2] && prim[2] <= b[5]);
      },
      no_pref,
      [&](int32_t pos) {
        const int32_t j = a.leaf_ids[pos];
        if ((int64_t)j <= i || (int64_t)j >= a.n) return;
        const double* tj = a.tri + 9 * (int64_t)j;
        if (tri_tri_intersect(a0, a1, a2, load_v3(tj), load_v3(tj + 3), load_v3(tj + 6), false,
                              a.eps))
          hit(j);
      });
  if (v < 0) atomicOr(a.status, 1);
}

// Walk 1: count every triangle's hits, keep the first kPairSlots.
__global__ __launch_bounds__(kBlock) void k_isect_count(IsectArgs a) {
  __shared__ int32_t s_stack[stack_lds_words(kBlock)];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  int32_t c = 0;
  isect_walk(a, i, lane_stack_base(s_stack), [&](int32_t j) {
    if (c < kPairSlots) a.slots[kPairSlots * i + c] = j;
    ++c;
  });
  a.npair[i] = c;
}

// Write the pairs at the scanned offsets; triangles with more hits than
// slots walk again (the walk is deterministic, so the same hits come in the
// same order).
__global__ __launch_bounds__(kBlock) void k_isect_emit(IsectArgs a) {
  __shared__ int32_t s_stack[stack_lds_words(kBlock)];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t c = (i < a.n) ? a.npair[i] : 0;
  if (c == 0) return;  // (no block-wide barrier follows)
  const int32_t o = a.poff[i];
  if (c <= kPairSlots) {
    for (int32_t k = 0; k < c; ++k) {
      a.first[o + k] = (int32_t)i;
      a.second[o + k] = a.slots[kPairSlots * i + k];
    }
    return;
  }
  int32_t m = 0;
  isect_walk(a, i, lane_stack_base(s_stack), [&](int32_t j) {
    if (m < c) {
      a.first[o + m] = (int32_t)i;
      a.second[o + m] = j;
    }
    ++m;
  });
}

__global__ __launch_bounds__(kBlock) void k_flag_emit(const int32_t* __restrict__ flag,
                                                      const int32_t* __restrict__ off, int64_t n,
                                                      int32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) out[off[i]] = (int32_t)i;
}

// Batched primal::intersect(Triangle3, Triangle3, includeBoundary, EPS).
__global__ __launch_bounds__(kBlock) void k_tri_tri(const double* __restrict__ t1,
                                                    const double* __restrict__ t2, int64_t n,
                                                    bool inc, double eps,
                                                    uint8_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* a = t1 + 9 * i;
  const double* b = t2 + 9 * i;
  out[i] = tri_tri_intersect(load_v3(a), load_v3(a + 3), load_v3(a + 6), load_v3(b), load_v3(b + 3),
                             load_v3(b + 6), inc, eps)
               ? 1
               : 0;
}

// Device allocations of one call, freed on every exit path.
// Bounding box of each cell of `npc` nodes (primal::compute_bounding_box of
// the cell's points: BoundingBox::addPoint of each, BoundingBox.hpp:466-482,
// a coordinate replaces the bound it is strictly below / above). A node index
// outside [0, nverts) sets bit 1 of *status.
__global__ __launch_bounds__(kBlock) void k_cell_boxes_n(const double* __restrict__ x,
                                                         const double* __restrict__ y,
                                                         const double* __restrict__ z,
                                                         int64_t nverts,
                                                         const int32_t* __restrict__ conn,
                                                         int64_t ncells, int npc,
                                                         double* __restrict__ boxes,
                                                         int32_t* __restrict__ status) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  double b[6] = {kDblMax, kDblMax, kDblMax, -kDblMax, -kDblMax, -kDblMax};
  for (int k = 0; k < npc; ++k) {
    const int32_t v = conn[c * npc + k];
    if (v < 0 || (int64_t)v >= nverts) {
      atomicOr(status, 1);
      continue;
    }
    const double p[3] = {x[v], y[v], z[v]};
    for (int d = 0; d < 3; ++d) {
      if (p[d] < b[d]) b[d] = p[d];
      if (p[d] > b[3 + d]) b[3 + d] = p[d];
    }
  }
  double* o = boxes + 6 * c;
  for (int k = 0; k < 6; ++k) o[k] = b[k];
}

// The first member of every candidate pair: query cell i at
// [offsets[i], offsets[i] + counts[i]).
__global__ __launch_bounds__(kBlock) void k_expand_first(const int32_t* __restrict__ offsets,
                                                         const int32_t* __restrict__ counts,
                                                         int64_t nq, int32_t* __restrict__ first) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const int32_t o = offsets[i], c = counts[i];
  for (int32_t j = 0; j < c; ++j) first[o + j] = (int32_t)i;
}

struct Scratch {
  std::vector<void*> ptrs;
  template <class T>
  int alloc(T** p, int64_t count) {
    *p = nullptr;
    const size_t bytes = sizeof(T) * (size_t)(count > 0 ? count : 1);
    if (hipMalloc((void**)p, bytes) != hipSuccess) {
      *p = nullptr;
      set_error("device allocation of %zu bytes failed", bytes);
      return AXSQ_ERR_NOMEM;
    }
    ptrs.push_back(*p);
    return AXSQ_OK;
  }
  void keep(void* p) {  // ownership passes to the caller
    for (auto& q : ptrs)
      if (q == p) q = nullptr;
  }
  ~Scratch() {
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
  }
};

}  // namespace
}  // namespace axsq

using namespace axsq;

extern "C" int axsq_tri_tri_intersect(const double* t1, const double* t2, int64_t n,
                                      int include_boundary, double eps, uint8_t* out,
                                      void* stream) {
  clear_error();
  if (n < 0 || (n > 0 && (!t1 || !t2 || !out))) {
    set_error("axsq_tri_tri_intersect: need n >= 0 and non-NULL arrays");
    return AXSQ_ERR_INVALID_ARG;
  }
  if (n == 0) return AXSQ_OK;
  hipStream_t s = (hipStream_t)stream;
  k_tri_tri<<<grid_for(n), kBlock, 0, s>>>(t1, t2, n, include_boundary != 0, eps, out);
  AXSQ_CHECK_HIP(hipGetLastError());
  AXSQ_CHECK_HIP(hipStreamSynchronize(s));
  return AXSQ_OK;
}

extern "C" int axsq_mesh_self_intersections(const double* x, const double* y, const double* z,
                                            int64_t nverts, const int32_t* tris, int64_t ntris,
                                            double threshold, int32_t** first, int32_t** second,
                                            int64_t* npairs, int32_t** degenerate,
                                            int64_t* ndegen, void* stream) {
  clear_error();
  if (!first || !second || !npairs || !degenerate || !ndegen) {
    set_error("axsq_mesh_self_intersections: NULL output");
    return AXSQ_ERR_INVALID_ARG;
  }
  *first = *second = *degenerate = nullptr;
  *npairs = *ndegen = 0;
  if (ntris < 0 || nverts < 0 || ntris > 0x3fffffffll || nverts > 0x7fffffffll ||
      (ntris > 0 && (!x || !y || !z || !tris || nverts < 1))) {
    set_error("axsq_mesh_self_intersections: invalid mesh arguments");
    return AXSQ_ERR_INVALID_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  Scratch out;  // the caller's arrays until they are handed over
  int32_t *pf = nullptr, *ps = nullptr, *pd = nullptr;
  if (ntris == 0) {
    AXSQ_TRY(out.alloc(&pf, 1));
    AXSQ_TRY(out.alloc(&ps, 1));
    AXSQ_TRY(out.alloc(&pd, 1));
  } else {
    const int64_t n = ntris;
    Workspace& ws = workspace();
    AXSQ_TRY(ws.mesh_tri.ensure(sizeof(double) * 9 * n));
    AXSQ_TRY(ws.mesh_box.ensure(sizeof(double) * 6 * n));
    AXSQ_TRY(ws.mesh_int.ensure(sizeof(int32_t) * (4 * n + 2) + 2 * sizeof(int64_t) + 64));
    AXSQ_TRY(ws.mesh_slots.ensure(sizeof(int32_t) * kPairSlots * n));
    double* tri = ws.mesh_tri.as<double>();
    double* boxes = ws.mesh_box.as<double>();
    int64_t* tot = reinterpret_cast<int64_t*>(ws.mesh_int.as<char>());
    int32_t* degf = reinterpret_cast<int32_t*>(tot + 2);
    int32_t* degoff = degf + n;
    int32_t* npair = degoff + n;
    int32_t* poff = npair + n;
    int32_t* status = poff + n;
    AXSQ_TRY(fill_i32(status, 0, 1, s));
    {
      ProfRange pr("mesh_isect_setup", s);
      k_tri_setup<<<grid_for(n), kBlock, 0, s>>>(x, y, z, tris, n, nverts, tri, boxes, degf,
                                                 status);
      AXSQ_CHECK_HIP(hipGetLastError());
    }
    int32_t st = 0;
    AXSQ_CHECK_HIP(hipMemcpyAsync(&st, status, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    AXSQ_CHECK_HIP(hipStreamSynchronize(s));
    if (st) {
      set_error("axsq_mesh_self_intersections: connectivity refers to a vertex outside [0, %lld)",
                (long long)nverts);
      return AXSQ_ERR_INVALID_ARG;
    }
    // the BVH over the corner boxes, default scale (MeshTester_detail.hpp:322-338)
    if (!ws.mesh_bvh) {
      ws.mesh_bvh = new axsq_bvh();
      hipGetDevice(&ws.mesh_bvh->device);
    }
    const int rc = bvh_build_into(ws.mesh_bvh, boxes, n, 1.000123, AXSQ_FP64, s);
    if (rc != AXSQ_OK) {
      axsq_bvh_destroy(ws.mesh_bvh);
      ws.mesh_bvh = nullptr;
      return rc;
    }
    const axsq_bvh* b = ws.mesh_bvh;
    IsectArgs a;
    a.nodes = b->nodes;
    a.ninner = (int32_t)(b->n - 1);
    a.leaf_ids = b->leaf_ids;
    a.tri = tri;
    a.boxes = boxes;
    a.n = n;
    a.eps = threshold;
    a.npair = npair;
    a.slots = ws.mesh_slots.as<int32_t>();
    a.poff = poff;
    a.status = b->status;
    {
      ProfRange pr("mesh_isect_walk", s);
      k_isect_count<<<grid_for(n), kBlock, 0, s>>>(a);
      AXSQ_CHECK_HIP(hipGetLastError());
    }
    AXSQ_TRY(scan_exclusive_i32(npair, poff, n, tot, ws, s));
    AXSQ_TRY(scan_exclusive_i32(degf, degoff, n, tot + 1, ws, s));
    int64_t totals[2] = {0, 0};
    int32_t walk_status = 0;
    AXSQ_CHECK_HIP(hipMemcpyAsync(totals, tot, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    AXSQ_CHECK_HIP(
        hipMemcpyAsync(&walk_status, b->status, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    AXSQ_CHECK_HIP(hipStreamSynchronize(s));
    if (walk_status) {
      AXSQ_TRY(fill_i32(b->status, 0, 1, s));
      set_error("traversal stopped: stack deeper than %d entries or malformed tree", kStack);
      return AXSQ_ERR_STACK_OVERFLOW;
    }
    AXSQ_TRY(out.alloc(&pf, totals[0]));
    AXSQ_TRY(out.alloc(&ps, totals[0]));
    AXSQ_TRY(out.alloc(&pd, totals[1]));
    a.first = pf;
    a.second = ps;
    {
      ProfRange pr("mesh_isect_emit", s);
      k_isect_emit<<<grid_for(n), kBlock, 0, s>>>(a);
      AXSQ_CHECK_HIP(hipGetLastError());
      k_flag_emit<<<grid_for(n), kBlock, 0, s>>>(degf, degoff, n, pd);
      AXSQ_CHECK_HIP(hipGetLastError());
    }
    AXSQ_CHECK_HIP(hipStreamSynchronize(s));
    prof_flush();
    *npairs = totals[0];
    *ndegen = totals[1];
  }
  out.keep(pf);
  out.keep(ps);
  out.keep(pd);
  *first = pf;
  *second = ps;
  *degenerate = pd;
  return AXSQ_OK;
}

extern "C" int axsq_mesh_candidates(const double* ix, const double* iy, const double* iz,
                                    int64_t inverts, const int32_t* icells, int64_t nicells,
                                    const double* qx, const double* qy, const double* qz,
                                    int64_t qnverts, const int32_t* qcells, int64_t nqcells,
                                    int nodes_per_cell, int32_t** first, int32_t** second,
                                    int64_t* npairs, void* stream) {
  clear_error();
  if (!first || !second || !npairs) {
    set_error("axsq_mesh_candidates: NULL output");
    return AXSQ_ERR_INVALID_ARG;
  }
  *first = *second = nullptr;
  *npairs = 0;
  if (nodes_per_cell < 1 || nodes_per_cell > 8) {
    set_error("axsq_mesh_candidates: nodes_per_cell %d outside [1, 8]", nodes_per_cell);
    return AXSQ_ERR_INVALID_ARG;
  }
  if (nicells < 0 || nqcells < 0 || nicells > 0x3fffffffll || nqcells > 0x3fffffffll ||
      inverts < 0 || qnverts < 0 || inverts > 0x7fffffffll || qnverts > 0x7fffffffll ||
      (nicells > 0 && (!ix || !iy || !iz || !icells)) ||
      (nqcells > 0 && (!qx || !qy || !qz || !qcells))) {
    set_error("axsq_mesh_candidates: invalid mesh arguments");
    return AXSQ_ERR_INVALID_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  Scratch out;
  int32_t *pf = nullptr, *ps = nullptr;
  int64_t total = 0;
  if (nicells == 0 || nqcells == 0) {
    AXSQ_TRY(out.alloc(&pf, 1));
    AXSQ_TRY(out.alloc(&ps, 1));
  } else {
    const int64_t ni = nicells, nq = nqcells;
    Workspace& ws = workspace();
    AXSQ_TRY(ws.mesh_box.ensure(sizeof(double) * 6 * ni));
    AXSQ_TRY(ws.mesh_tri.ensure(sizeof(double) * 6 * nq));
    // mesh_int: totals (2 x int64), status, then offsets and counts
    AXSQ_TRY(ws.mesh_int.ensure(2 * sizeof(int64_t) + sizeof(int32_t) * (2 * nq + 2)));
    double* ibox = ws.mesh_box.as<double>();
    double* qbox = ws.mesh_tri.as<double>();
    int64_t* tot = reinterpret_cast<int64_t*>(ws.mesh_int.as<char>());
    int32_t* status = reinterpret_cast<int32_t*>(tot + 2);
    int32_t* offsets = status + 2;
    int32_t* counts = offsets + nq;
    AXSQ_TRY(fill_i32(status, 0, 1, s));
    k_cell_boxes_n<<<grid_for(ni), kBlock, 0, s>>>(ix, iy, iz, inverts, icells, ni,
                                                   nodes_per_cell, ibox, status);
    AXSQ_CHECK_HIP(hipGetLastError());
    k_cell_boxes_n<<<grid_for(nq), kBlock, 0, s>>>(qx, qy, qz, qnverts, qcells, nq,
                                                   nodes_per_cell, qbox, status);
    AXSQ_CHECK_HIP(hipGetLastError());
    int32_t st = 0;
    AXSQ_CHECK_HIP(hipMemcpyAsync(&st, status, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    AXSQ_CHECK_HIP(hipStreamSynchronize(s));
    if (st) {
      set_error("axsq_mesh_candidates: a cell refers to a vertex outside its mesh");
      return AXSQ_ERR_INVALID_ARG;
    }
    // the BVH over the insert mesh's boxes, default scale
    // (quest_candidates_example.cpp:420-423)
    if (!ws.mesh_bvh) {
      ws.mesh_bvh = new axsq_bvh();
      hipGetDevice(&ws.mesh_bvh->device);
    }
    const int rc = bvh_build_into(ws.mesh_bvh, ibox, ni, 1.000123, AXSQ_FP64, s);
    if (rc != AXSQ_OK) {
      axsq_bvh_destroy(ws.mesh_bvh);
      ws.mesh_bvh = nullptr;
      return rc;
    }
    const axsq_bvh* b = ws.mesh_bvh;
    // findBoundingBoxes for every query box (:426-434): one walk, the first
    // ids kept, the rest walked again
    const int kstash = stash_slots(nq);
    AXSQ_TRY(ws.find_stash.ensure(sizeof(int32_t) * kstash * (size_t)nq));
    AXSQ_TRY(ws.find_cnt.ensure(sizeof(int32_t) * (size_t)nq));
    AXSQ_TRY(ws.find_over.ensure(sizeof(int32_t) * (size_t)nq));
    AXSQ_TRY(ws.find_n.ensure(sizeof(int64_t) * 2));
    uint32_t* nover = reinterpret_cast<uint32_t*>(ws.find_n.as<int64_t>());
    AXSQ_TRY(fill_i32(reinterpret_cast<int32_t*>(nover), 0, 1, s));
    const double tol = 2.220446049250313e-16;  // BVH default tolerance (unused by boxes)
    AXSQ_TRY(find_launch(b, AXSQ_QUERY_BOXES, 0, qbox, nullptr, nullptr, 6, nq, tol, counts,
                         nullptr, nullptr, nullptr, ws.find_stash.as<int32_t>(), kstash,
                         ws.find_cnt.as<int32_t>(), ws.find_over.as<int32_t>(), nover, s));
    AXSQ_TRY(scan_exclusive_i32(counts, offsets, nq, tot, ws, s));
    int64_t tn[2] = {0, 0};
    int32_t walk_status = 0;
    AXSQ_CHECK_HIP(hipMemcpyAsync(&tn[0], tot, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    AXSQ_CHECK_HIP(hipMemcpyAsync(&tn[1], nover, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    AXSQ_CHECK_HIP(
        hipMemcpyAsync(&walk_status, b->status, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    AXSQ_CHECK_HIP(hipStreamSynchronize(s));
    if (walk_status) {
      AXSQ_TRY(fill_i32(b->status, 0, 1, s));
      set_error("traversal stopped: stack deeper than %d entries or malformed tree", kStack);
      return AXSQ_ERR_STACK_OVERFLOW;
    }
    total = tn[0];
    if (total > 0x7fffffffll) {
      set_error("axsq_mesh_candidates: %lld candidate pairs exceed the int32 index range",
                (long long)total);
      return AXSQ_ERR_INVALID_ARG;
    }
    const int64_t n_over = (int64_t)(uint32_t)tn[1];
    AXSQ_TRY(out.alloc(&pf, total));
    AXSQ_TRY(out.alloc(&ps, total));
    // the pair list (:440-458): query cell, then its candidates in order
    AXSQ_TRY(stash_copy_launch(ws.find_stash.as<int32_t>(), kstash, ws.find_cnt.as<int32_t>(),
                               offsets, nq, ps, s));
    if (n_over > 0)
      AXSQ_TRY(find_launch(b, AXSQ_QUERY_BOXES, 2, qbox, nullptr, nullptr, 6, n_over, tol,
                           nullptr, counts, offsets, ps, nullptr, 0, nullptr,
                           ws.find_over.as<int32_t>(), nullptr, s));
    k_expand_first<<<grid_for(nq), kBlock, 0, s>>>(offsets, counts, nq, pf);
    AXSQ_CHECK_HIP(hipGetLastError());
    AXSQ_CHECK_HIP(hipStreamSynchronize(s));
    prof_flush();
  }
  out.keep(pf);
  out.keep(ps);
  *first = pf;
  *second = ps;
  *npairs = total;
  return AXSQ_OK;
}
