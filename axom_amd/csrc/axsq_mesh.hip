// axsq_mesh.hip -- triangle-triangle intersection and the BVH mesh
// self-intersection test (quest::findTriMeshIntersectionsBVH) for gfx950.
//
// The pipeline of quest/detail/MeshTester_detail.hpp:159-341 -- per-triangle
// corner boxes and degenerate flags, a BVH over the boxes, findBoundingBoxes
// of every box, the narrow phase on the pairs i < j -- with the sequential
// path's output order: pairs grouped by the first triangle, the second in
// the first's candidate (BVH traversal) order; degenerate indices ascending.
// The narrow phase runs one triangle per lane over its candidate list, in two
// passes (flags + counts, then emit after a scan), so the order needs no
// atomics.
#include <hip/hip_runtime.h>

#include "axsq_internal.hpp"
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

// Narrow phase, pass 1: for triangle i, flag each candidate j > i that
// intersects it (boundary excluded, tolerance eps) and count the flags.
__global__ __launch_bounds__(kBlock) void k_tt_flags(const double* __restrict__ tri,
                                                     const int32_t* __restrict__ offsets,
                                                     const int32_t* __restrict__ counts,
                                                     const int32_t* __restrict__ cands, int64_t n,
                                                     double eps, uint8_t* __restrict__ flag,
                                                     int32_t* __restrict__ npair) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* ti = tri + 9 * i;
  const V3 a0 = load_v3(ti), a1 = load_v3(ti + 3), a2 = load_v3(ti + 6);
  const int32_t off = offsets[i], cnt = counts[i];
  int32_t c = 0;
  for (int32_t k = 0; k < cnt; ++k) {
    const int32_t j = cands[off + k];
    bool f = false;
    if ((int64_t)j > i && j < n) {
      const double* tj = tri + 9 * (int64_t)j;
      f = tri_tri_intersect(a0, a1, a2, load_v3(tj), load_v3(tj + 3), load_v3(tj + 6), false, eps);
    }
    flag[off + k] = f ? 1 : 0;
    c += f ? 1 : 0;
  }
  npair[i] = c;
}

// Narrow phase, pass 2: write the flagged pairs at the scanned offsets.
__global__ __launch_bounds__(kBlock) void k_tt_emit(const int32_t* __restrict__ offsets,
                                                    const int32_t* __restrict__ counts,
                                                    const int32_t* __restrict__ cands,
                                                    const uint8_t* __restrict__ flag,
                                                    const int32_t* __restrict__ poff, int64_t n,
                                                    int32_t* __restrict__ first,
                                                    int32_t* __restrict__ second) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t off = offsets[i], cnt = counts[i];
  int32_t o = poff[i];
  for (int32_t k = 0; k < cnt; ++k) {
    if (flag[off + k]) {
      first[o] = (int32_t)i;
      second[o] = cands[off + k];
      ++o;
    }
  }
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

struct BvhGuard {
  axsq_bvh* b = nullptr;
  ~BvhGuard() {
    if (b) axsq_bvh_destroy(b);
  }
};

int read_total(const int64_t* dev, int64_t* host, hipStream_t s) {
  AXSQ_CHECK_HIP(hipMemcpyAsync(host, dev, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  AXSQ_CHECK_HIP(hipStreamSynchronize(s));
  return AXSQ_OK;
}

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
  Scratch sc;
  int32_t *pf = nullptr, *ps = nullptr, *pd = nullptr;
  if (ntris == 0) {
    AXSQ_TRY(sc.alloc(&pf, 1));
    AXSQ_TRY(sc.alloc(&ps, 1));
    AXSQ_TRY(sc.alloc(&pd, 1));
    sc.keep(pf);
    sc.keep(ps);
    sc.keep(pd);
    *first = pf;
    *second = ps;
    *degenerate = pd;
    return AXSQ_OK;
  }
  const int64_t n = ntris;
  double *tri = nullptr, *boxes = nullptr;
  int32_t *degf = nullptr, *degoff = nullptr, *counts = nullptr, *offsets = nullptr;
  int32_t *npair = nullptr, *poff = nullptr, *status = nullptr;
  int64_t* tot = nullptr;
  AXSQ_TRY(sc.alloc(&tri, 9 * n));
  AXSQ_TRY(sc.alloc(&boxes, 6 * n));
  AXSQ_TRY(sc.alloc(&degf, n));
  AXSQ_TRY(sc.alloc(&degoff, n));
  AXSQ_TRY(sc.alloc(&counts, n));
  AXSQ_TRY(sc.alloc(&offsets, n));
  AXSQ_TRY(sc.alloc(&npair, n));
  AXSQ_TRY(sc.alloc(&poff, n));
  AXSQ_TRY(sc.alloc(&status, 1));
  AXSQ_TRY(sc.alloc(&tot, 2));
  AXSQ_TRY(fill_i32(status, 0, 1, s));
  {
    ProfRange pr("mesh_isect_setup", s);
    k_tri_setup<<<grid_for(n), kBlock, 0, s>>>(x, y, z, tris, n, nverts, tri, boxes, degf, status);
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
  // broad phase: findBoundingBoxes of every triangle box against the BVH of
  // all of them (MeshTester_detail.hpp:322-338)
  BvhGuard bg;
  AXSQ_TRY(axsq_bvh_build(&bg.b, boxes, n, 1.000123, AXSQ_FP64, stream));
  int64_t total = 0;
  AXSQ_TRY(axsq_bvh_find_count(bg.b, AXSQ_QUERY_BOXES, boxes, nullptr, nullptr, 6, n, 0.0, offsets,
                               counts, &total, stream));
  int32_t* cands = nullptr;
  uint8_t* flag = nullptr;
  AXSQ_TRY(sc.alloc(&cands, total));
  AXSQ_TRY(sc.alloc(&flag, total));
  if (total > 0)
    AXSQ_TRY(axsq_bvh_find_fill(bg.b, AXSQ_QUERY_BOXES, boxes, nullptr, nullptr, 6, n, 0.0,
                                offsets, cands, stream));
  // narrow phase (MeshTester_detail.hpp:235-289)
  Workspace& ws = workspace();
  {
    ProfRange pr("mesh_isect_narrow", s);
    k_tt_flags<<<grid_for(n), kBlock, 0, s>>>(tri, offsets, counts, cands, n, threshold, flag,
                                              npair);
    AXSQ_CHECK_HIP(hipGetLastError());
  }
  AXSQ_TRY(scan_exclusive_i32(npair, poff, n, tot, ws, s));
  AXSQ_TRY(scan_exclusive_i32(degf, degoff, n, tot + 1, ws, s));
  int64_t totals[2] = {0, 0};
  AXSQ_CHECK_HIP(hipMemcpyAsync(totals, tot, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  AXSQ_CHECK_HIP(hipStreamSynchronize(s));
  AXSQ_TRY(sc.alloc(&pf, totals[0]));
  AXSQ_TRY(sc.alloc(&ps, totals[0]));
  AXSQ_TRY(sc.alloc(&pd, totals[1]));
  {
    ProfRange pr("mesh_isect_emit", s);
    k_tt_emit<<<grid_for(n), kBlock, 0, s>>>(offsets, counts, cands, flag, poff, n, pf, ps);
    AXSQ_CHECK_HIP(hipGetLastError());
    k_flag_emit<<<grid_for(n), kBlock, 0, s>>>(degf, degoff, n, pd);
    AXSQ_CHECK_HIP(hipGetLastError());
  }
  AXSQ_CHECK_HIP(hipStreamSynchronize(s));
  prof_flush();
  sc.keep(pf);
  sc.keep(ps);
  sc.keep(pd);
  *first = pf;
  *second = ps;
  *degenerate = pd;
  *npairs = totals[0];
  *ndegen = totals[1];
  return AXSQ_OK;
}
