// axsq_device.hpp -- device geometry for the gfx950 spatial-query kernels.
//
// Every function performs the same double operations, in the same order, as
// the reference's primal/spin code it cites (paths relative to the
// reference's src/axom/), so results round identically. The library is
// compiled with -ffp-contract=off so no multiply-add is fused.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace axsq {

constexpr int32_t kBarrier = -2000000000;  // bvh_traverse.hpp:80
constexpr int kStack = 64;                 // bvh_traverse.hpp:79
constexpr double kDblMax = 1.7976931348623157e308;

struct Box {
  double lo[3];
  double hi[3];
};

// FP64 inner node: both child boxes and both child links in one 128-byte
// cache line. Link >= 0: inner node index; < 0: -(leaf position + 1).
// flags bit 0 / bit 1: left / right box is valid (BoundingBox::isValid of the
// stored box, precomputed so traversal does not re-test it). [lo, hi] is the
// range of leaf positions under the node (Karras 2012: the node's index is lo
// or hi, and the inner nodes of its subtree are the hi - lo indices starting
// at lo, or at lo + 1 when the index is hi); the left child covers [lo, mid],
// the right child [mid + 1, hi].
struct __attribute__((aligned(128))) Node64 {
  double l[6];  // left  min xyz, max xyz
  double r[6];  // right min xyz, max xyz
  int32_t lc, rc;
  int32_t flags;
  int32_t lo, hi, mid;
  int32_t pad[2];
};
static_assert(sizeof(Node64) == 128, "Node64 must be one cache line");

// FP32 inner node (mixed precision): boxes rounded outward, 64 bytes.
struct __attribute__((aligned(64))) Node32 {
  float l[6];
  float r[6];
  int32_t lc, rc;
  int32_t flags;  // as Node64::flags, for the FP32 boxes
  int32_t pad;
};
static_assert(sizeof(Node32) == 64, "Node32 must be half a cache line");

// Triangle record in BVH leaf order: A, B, C (as in the connectivity) and
// the original cell id. 80 bytes.
struct __attribute__((aligned(16))) TriRec {
  double v[9];
  int32_t id;
  int32_t pad;
};
static_assert(sizeof(TriRec) == 80, "TriRec is 80 bytes");

// core/utilities/Utilities.hpp:43-46,318-321
__device__ __forceinline__ bool near_eq(double a, double b, double thresh) {
  double d = a - b;
  d = (d < 0.0) ? -d : d;
  return d <= thresh;
}
// primal/operators/detail/fuzzy_comparators.hpp:26-44
__device__ __forceinline__ bool is_gt(double x, double y, double e) { return (x > y) && !near_eq(x, y, e); }
__device__ __forceinline__ bool is_lt(double x, double y, double e) { return (x < y) && !near_eq(x, y, e); }
__device__ __forceinline__ bool is_leq(double x, double y, double e) { return !is_gt(x, y, e); }
__device__ __forceinline__ bool is_geq(double x, double y, double e) { return !is_lt(x, y, e); }

// primal/geometry/Vector.hpp:543-552: the sum starts from +0.0, which turns a
// leading -0.0 product into +0.0 exactly as the reference does.
__device__ __forceinline__ double dot3(double a0, double a1, double a2, double b0, double b1,
                                       double b2) {
  double r = 0.0;
  r += a0 * b0;
  r += a1 * b1;
  r += a2 * b2;
  return r;
}

// primal/geometry/Vector.hpp:564-571 (Determinants.hpp:31-38)
__device__ __forceinline__ void cross3(const double u[3], const double v[3], double o[3]) {
  o[0] = u[1] * v[2] - v[1] * u[2];
  o[1] = v[0] * u[2] - u[0] * v[2];
  o[2] = u[0] * v[1] - v[0] * u[1];
}

// Vector::unitVector, Vector.hpp:477-490; v /= s multiplies by 1/s
// (core/NumericArray.hpp:510-514).
__device__ __forceinline__ void unit3(const double v[3], double o[3]) {
  const double len_sq = dot3(v[0], v[1], v[2], v[0], v[1], v[2]);
  if (len_sq >= 1e-50) {
    const double inv = 1. / sqrt(len_sq);
    o[0] = v[0] * inv;
    o[1] = v[1] * inv;
    o[2] = v[2] * inv;
  } else {
    o[0] = 1.0;
    o[1] = 0.0;
    o[2] = 0.0;
  }
}

// BoundingBox::isValid, BoundingBox.hpp:451-462
__device__ __forceinline__ bool box_valid(const double b[6]) {
  return !(b[0] > b[3]) && !(b[1] > b[4]) && !(b[2] > b[5]);
}
__device__ __forceinline__ bool box_valid_f(const float b[6]) {
  return !(b[0] > b[3]) && !(b[1] > b[4]) && !(b[2] > b[5]);
}

// BoundingBox::addBox, BoundingBox.hpp:486-508 (addPoint :466-484 twice)
__device__ __forceinline__ void box_add_box(double a[6], const double o[6]) {
  if (box_valid(a)) {
    if (box_valid(o)) {
      for (int pass = 0; pass < 2; ++pass) {
        const double* p = o + 3 * pass;
        for (int d = 0; d < 3; ++d) {
          if (p[d] < a[d]) a[d] = p[d];
          if (p[d] > a[3 + d]) a[3 + d] = p[d];
        }
      }
    }
  } else {
    for (int k = 0; k < 6; ++k) a[k] = o[k];
  }
}

// BoundingBox::scale, BoundingBox.hpp:547-561 (+ checkAndFixBounds :575-585)
__device__ __forceinline__ void box_scale(double b[6], double sf) {
  if (!box_valid(b)) return;
  const double s = sf * 0.5;
  double mid[3], r[3];
  for (int d = 0; d < 3; ++d) {
    mid[d] = 0.5 * (b[d] + b[3 + d]);
    r[d] = (b[3 + d] - b[d]) * s;
  }
  for (int d = 0; d < 3; ++d) {
    double lo = mid[d] - r[d];
    double hi = mid[d] + r[d];
    if (lo > hi) {
      const double t = lo;
      lo = hi;
      hi = t;
    }
    b[d] = lo;
    b[3 + d] = hi;
  }
}

// 10-bit spread for the 30-bit Morton code of spin/MortonIndex.hpp:369-470.
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// morton32_encode of the normalised centroid, build_radix_tree.hpp:48-65,154-175
__device__ __forceinline__ uint32_t morton_of(const double c[3], const double minc[3],
                                              const double inv[3]) {
  uint32_t code = 0;
  for (int d = 0; d < 3; ++d) {
    const double u = (c[d] - minc[d]) * inv[d];
    const int32_t q = (int32_t)fmin(fmax(u * 1024.0, 0.0), 1023.0);
    code |= spread10((uint32_t)q) << d;
  }
  return code;
}

// 16-bit spread and the 32-bit 2-D Morton code (morton32_encode with Dims = 2:
// 16 bits per axis, x in the even bits; build_radix_tree.hpp:48-65 and the
// 2-D Mortonizer, spin/MortonIndex.hpp:226-318, whose int32 expansion runs
// the shifts 16, 8, 4, 2, 1).
__device__ __forceinline__ uint32_t spread16(uint32_t v) {
  v &= 0xffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}
__device__ __forceinline__ uint32_t morton2_of(const double c[3], const double minc[3],
                                               const double inv[3]) {
  uint32_t code = 0;
  for (int d = 0; d < 2; ++d) {
    const double u = (c[d] - minc[d]) * inv[d];
    const int32_t q = (int32_t)fmin(fmax(u * 65536.0, 0.0), 65535.0);
    code |= spread16((uint32_t)q) << d;
  }
  return code;
}

// Karras delta with index tie-break, build_radix_tree.hpp:265-287
__device__ __forceinline__ int32_t karras_delta(int32_t a, int32_t b, int32_t inner,
                                                const uint32_t* __restrict__ mc) {
  const bool oor = (b < 0 || b > inner);
  const int32_t bb = oor ? 0 : b;
  uint32_t ex = mc[a] ^ mc[bb];
  const bool tie = (ex == 0);
  if (tie) ex = (uint32_t)a ^ (uint32_t)bb;
  int32_t cnt = ex ? __clz((int)ex) : 32;
  if (tie) cnt += 32;
  return oor ? -1 : cnt;
}

// squared_distance(Point, Box), primal/operators/squared_distance.hpp:77-102.
// For a point inside the box the clamped vector is exactly zero, so the
// contains() shortcut of the reference yields the same 0. The clamped
// difference (clampVal, core/utilities/Utilities.hpp:147-150) is taken as
// the gap max(lo - q, q - hi, 0): on a valid box at most one of lo - q and
// q - hi is positive, and it equals the clamped difference up to sign, which
// the squaring erases, so the sum is bit-identical. Max/min of subtraction
// results need no canonicalisation of loaded operands, which the clamp
// form costs once per box coordinate.
// A NaN coordinate gets gap 0 (max drops NaN), as it passes contains()
// (BoundingBox.hpp:390-401) in that dimension. The reference's distance is
// then 0 when the other coordinates lie in the box and NaN otherwise; the
// walks reproduce its pruning by starting such a query at bound 0 (see
// query_has_nan), so that only a sum of 0 passes.
__device__ __forceinline__ double sqdist_box(double q0, double q1, double q2, const double b[6]) {
  const double e0 = fmax(fmax(b[0] - q0, q0 - b[3]), 0.0);
  const double e1 = fmax(fmax(b[1] - q1, q1 - b[4]), 0.0);
  const double e2 = fmax(fmax(b[2] - q2, q2 - b[5]), 0.0);
  return dot3(e0, e1, e2, e0, e1, e2);
}

// A query with a NaN coordinate: every triangle distance is NaN, so no
// candidate is taken and the running minimum of the reference stays at
// DBL_MAX, against which a box passes exactly when its distance is 0 (see
// sqdist_box). Its walk therefore starts at bound 0 in place of DBL_MAX.
__device__ __forceinline__ bool query_has_nan(const double q[3]) {
  return __builtin_isnan(q[0]) || __builtin_isnan(q[1]) || __builtin_isnan(q[2]);
}

// closest_point(Point, Triangle, loc, EPS), primal/operators/closest_point.hpp:162-290.
// Location codes: 0,1,2 vertex A,B,C; -1 edge AB; -3 edge AC; -2 edge BC; 3 face.
// Also returns the barycentric weights of the result (for the adjoint).
template <bool WANT_W>
__device__ __forceinline__ int closest_point(const double P[3], const double* __restrict__ T,
                                             double EPS, double cp[3], double w[3]) {
  const double* A = T;
  const double* B = T + 3;
  const double* C = T + 6;
  const double ab0 = B[0] - A[0], ab1 = B[1] - A[1], ab2 = B[2] - A[2];
  const double ac0 = C[0] - A[0], ac1 = C[1] - A[1], ac2 = C[2] - A[2];
  const double ap0 = P[0] - A[0], ap1 = P[1] - A[1], ap2 = P[2] - A[2];
  const double d1 = dot3(ab0, ab1, ab2, ap0, ap1, ap2);
  const double d2 = dot3(ac0, ac1, ac2, ap0, ap1, ap2);
  if (is_leq(d1, 0.0, EPS) && is_leq(d2, 0.0, EPS)) {
    cp[0] = A[0]; cp[1] = A[1]; cp[2] = A[2];
    if (WANT_W) { w[0] = 1.0; w[1] = 0.0; w[2] = 0.0; }
    return 0;
  }
  const double bp0 = P[0] - B[0], bp1 = P[1] - B[1], bp2 = P[2] - B[2];
  const double d3 = dot3(ab0, ab1, ab2, bp0, bp1, bp2);
  const double d4 = dot3(ac0, ac1, ac2, bp0, bp1, bp2);
  if (is_geq(d3, 0.0, EPS) && is_leq(d4, d3, EPS)) {
    cp[0] = B[0]; cp[1] = B[1]; cp[2] = B[2];
    if (WANT_W) { w[0] = 0.0; w[1] = 1.0; w[2] = 0.0; }
    return 1;
  }
  const double vc = d1 * d4 - d3 * d2;
  if (is_leq(vc, 0.0, EPS) && is_geq(d1, 0.0, EPS) && is_leq(d3, 0.0, EPS) &&
      !near_eq(d1, d3, EPS)) {
    const double v = d1 / (d1 - d3);
    cp[0] = A[0] + ab0 * v; cp[1] = A[1] + ab1 * v; cp[2] = A[2] + ab2 * v;
    if (WANT_W) { w[0] = 1.0 - v; w[1] = v; w[2] = 0.0; }
    return -1;
  }
  const double cq0 = P[0] - C[0], cq1 = P[1] - C[1], cq2 = P[2] - C[2];
  const double d5 = dot3(ab0, ab1, ab2, cq0, cq1, cq2);
  const double d6 = dot3(ac0, ac1, ac2, cq0, cq1, cq2);
  if (is_geq(d6, 0.0, EPS) && is_leq(d5, d6, EPS)) {
    cp[0] = C[0]; cp[1] = C[1]; cp[2] = C[2];
    if (WANT_W) { w[0] = 0.0; w[1] = 0.0; w[2] = 1.0; }
    return 2;
  }
  const double vb = d5 * d2 - d1 * d6;
  if (is_leq(vb, 0.0, EPS) && is_geq(d2, 0.0, EPS) && is_leq(d6, 0.0, EPS)) {
    const double ww = d2 / (d2 - d6);
    cp[0] = A[0] + ac0 * ww; cp[1] = A[1] + ac1 * ww; cp[2] = A[2] + ac2 * ww;
    if (WANT_W) { w[0] = 1.0 - ww; w[1] = 0.0; w[2] = ww; }
    return -3;
  }
  const double va = d3 * d6 - d5 * d4;
  if (is_leq(va, 0.0, EPS) && is_geq(d4 - d3, 0.0, EPS) && is_geq(d5 - d6, 0.0, EPS)) {
    const double ww = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    cp[0] = B[0] + (C[0] - B[0]) * ww;
    cp[1] = B[1] + (C[1] - B[1]) * ww;
    cp[2] = B[2] + (C[2] - B[2]) * ww;
    if (WANT_W) { w[0] = 0.0; w[1] = 1.0 - ww; w[2] = ww; }
    return -2;
  }
  const double denom = 1.0 / (va + vb + vc);
  const double v = vb * denom;
  const double ww = vc * denom;
  cp[0] = A[0] + (ab0 * v + ac0 * ww);
  cp[1] = A[1] + (ab1 * v + ac1 * ww);
  cp[2] = A[2] + (ab2 * v + ac2 * ww);
  if (WANT_W) { w[0] = 1.0 - v - ww; w[1] = v; w[2] = ww; }
  return 3;
}

// The same closest point, written for a lane that runs alone in its wave:
// every dot product and every candidate division is computed up front as
// independent chains (instruction-level parallelism instead of a serial
// chain of early exits), then the regions are tested in the reference's order
// and the first match is returned. Each returned value is computed by the
// same operations as in closest_point<>, so the results are bit-identical.
__device__ __forceinline__ int closest_point_ilp(const double P[3], const double* __restrict__ T,
                                                 double EPS, double cp[3]) {
  const double* A = T;
  const double* B = T + 3;
  const double* C = T + 6;
  const double ab0 = B[0] - A[0], ab1 = B[1] - A[1], ab2 = B[2] - A[2];
  const double ac0 = C[0] - A[0], ac1 = C[1] - A[1], ac2 = C[2] - A[2];
  const double ap0 = P[0] - A[0], ap1 = P[1] - A[1], ap2 = P[2] - A[2];
  const double bp0 = P[0] - B[0], bp1 = P[1] - B[1], bp2 = P[2] - B[2];
  const double cq0 = P[0] - C[0], cq1 = P[1] - C[1], cq2 = P[2] - C[2];
  const double d1 = dot3(ab0, ab1, ab2, ap0, ap1, ap2);
  const double d2 = dot3(ac0, ac1, ac2, ap0, ap1, ap2);
  const double d3 = dot3(ab0, ab1, ab2, bp0, bp1, bp2);
  const double d4 = dot3(ac0, ac1, ac2, bp0, bp1, bp2);
  const double d5 = dot3(ab0, ab1, ab2, cq0, cq1, cq2);
  const double d6 = dot3(ac0, ac1, ac2, cq0, cq1, cq2);
  const double vc = d1 * d4 - d3 * d2;
  const double vb = d5 * d2 - d1 * d6;
  const double va = d3 * d6 - d5 * d4;
  int loc;
  if (is_leq(d1, 0.0, EPS) && is_leq(d2, 0.0, EPS)) {
    loc = 0;
  } else if (is_geq(d3, 0.0, EPS) && is_leq(d4, d3, EPS)) {
    loc = 1;
  } else if (is_leq(vc, 0.0, EPS) && is_geq(d1, 0.0, EPS) && is_leq(d3, 0.0, EPS) &&
             !near_eq(d1, d3, EPS)) {
    loc = -1;
  } else if (is_geq(d6, 0.0, EPS) && is_leq(d5, d6, EPS)) {
    loc = 2;
  } else if (is_leq(vb, 0.0, EPS) && is_geq(d2, 0.0, EPS) && is_leq(d6, 0.0, EPS)) {
    loc = -3;
  } else if (is_leq(va, 0.0, EPS) && is_geq(d4 - d3, 0.0, EPS) && is_geq(d5 - d6, 0.0, EPS)) {
    loc = -2;
  } else {
    loc = 3;
  }
  // The one division the region needs, with its own operands: the edge
  // parameters d1 / (d1 - d3), d2 / (d2 - d6), (d4 - d3) / ((d4 - d3) +
  // (d5 - d6)), or 1 / (va + vb + vc) for the face (the vertex regions use
  // none). One division per lane instead of four, with no branch.
  double num = 1.0, den = va + vb + vc;
  if (loc == -1) {
    num = d1;
    den = d1 - d3;
  } else if (loc == -3) {
    num = d2;
    den = d2 - d6;
  } else if (loc == -2) {
    num = d4 - d3;
    den = (d4 - d3) + (d5 - d6);
  }
  const double quo = num / den;
  switch (loc) {
    case 0:
      cp[0] = A[0]; cp[1] = A[1]; cp[2] = A[2];
      break;
    case 1:
      cp[0] = B[0]; cp[1] = B[1]; cp[2] = B[2];
      break;
    case 2:
      cp[0] = C[0]; cp[1] = C[1]; cp[2] = C[2];
      break;
    case -1:
      cp[0] = A[0] + ab0 * quo; cp[1] = A[1] + ab1 * quo; cp[2] = A[2] + ab2 * quo;
      break;
    case -3:
      cp[0] = A[0] + ac0 * quo; cp[1] = A[1] + ac1 * quo; cp[2] = A[2] + ac2 * quo;
      break;
    case -2:
      cp[0] = B[0] + (C[0] - B[0]) * quo;
      cp[1] = B[1] + (C[1] - B[1]) * quo;
      cp[2] = B[2] + (C[2] - B[2]) * quo;
      break;
    default: {
      const double v = vb * quo;
      const double ww = vc * quo;
      cp[0] = A[0] + (ab0 * v + ac0 * ww);
      cp[1] = A[1] + (ab1 * v + ac1 * ww);
      cp[2] = A[2] + (ab2 * v + ac2 * ww);
    }
  }
  return loc;
}

// Triangle::normal, primal/geometry/Triangle.hpp:99-102
__device__ __forceinline__ void tri_normal(const double* T, double n[3]) {
  const double u[3] = {T[3] - T[0], T[4] - T[1], T[5] - T[2]};
  const double v[3] = {T[6] - T[0], T[7] - T[1], T[8] - T[2]};
  cross3(u, v, n);
}

// Triangle::angle, Triangle.hpp:385-400. The three corners are picked with
// selects, not by a runtime index, so the triangle stays in registers.
__device__ __forceinline__ double tri_angle(const double* T, int idx) {
  double pt[3], p1[3], p2[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double A = T[d], B = T[3 + d], C = T[6 + d];
    pt[d] = (idx == 0) ? A : (idx == 1) ? B : C;
    p1[d] = (idx == 0) ? B : (idx == 1) ? C : A;
    p2[d] = (idx == 0) ? C : (idx == 1) ? A : B;
  }
  const double a[3] = {p1[0] - pt[0], p1[1] - pt[1], p1[2] - pt[2]};
  const double b[3] = {p2[0] - pt[0], p2[1] - pt[1], p2[2] - pt[2]};
  double ua[3], ub[3];
  unit3(a, ua);
  unit3(b, ub);
  double dp = dot3(ua[0], ua[1], ua[2], ub[0], ub[1], ub[2]);
  dp = (dp < -1.0) ? -1.0 : (dp > 1.0) ? 1.0 : dp;
  return acos(dp);
}

// intersect_ray_bbox_test per axis, primal/operators/detail/intersect_ray_impl.hpp:150-187,
// with utilities::max/min (core/utilities/Utilities.hpp:80-96).
__device__ __forceinline__ bool ray_slab(double x0, double n, double bmin, double bmax,
                                         double& tmin, double& tmax, double eps) {
  if (near_eq(n, 0.0, eps)) return !((x0 < bmin) || (x0 > bmax));
  const double invn = 1.0 / n;
  double t1 = (bmin - x0) * invn;
  double t2 = (bmax - x0) * invn;
  if (t1 > t2) {
    const double t = t1;
    t1 = t2;
    t2 = t;
  }
  tmin = (t1 < tmin) ? tmin : t1;
  tmax = (t2 < tmax) ? t2 : tmax;
  return !(tmin > tmax);
}

// intersect_ray(Ray, Box, ip, EPS), intersect_ray_impl.hpp:321-351
__device__ __forceinline__ bool ray_box(const double r[6], const double b[6], double eps) {
  double tmin = 2.2250738585072014e-308;  // DBL_MIN: a ray, not a line
  double tmax = kDblMax;
  bool hit = ray_slab(r[0], r[3], b[0], b[3], tmin, tmax, eps);
  hit = hit && ray_slab(r[1], r[4], b[1], b[4], tmin, tmax, eps);
  hit = hit && ray_slab(r[2], r[5], b[2], b[5], tmin, tmax, eps);
  return hit;
}

// ---- FloatType = float (spin::BVH<NDIMS, Exec, float>) ---------------------
// A float tree keeps its boxes as doubles holding float values (exact), so
// unions, comparisons and the point / box predicates give the float results
// unchanged; the steps below round to float where the reference's float
// arithmetic does.
constexpr float kFltMax = 3.40282347e38f;

// BoundingBox<float>::scale (BoundingBox.hpp:547-561): the centroid
// float(0.5 * (min + max)) with the sum in float, the radius
// float(range * float(sf * 0.5)) (the product of two floats is exact in
// double), then min/max in float and checkAndFixBounds.
__device__ __forceinline__ void box_scale_f32(double b[6], float sf) {
  if (!box_valid(b)) return;
  const float s = (float)((double)sf * 0.5);
  float lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    const float mn = (float)b[d], mx = (float)b[3 + d];
    const float mid = (float)(0.5 * (double)(mn + mx));
    const float r = (float)((double)(mx - mn) * (double)s);
    lo[d] = mid - r;
    hi[d] = mid + r;
  }
  for (int d = 0; d < 3; ++d) {
    float l = lo[d], h = hi[d];
    if (l > h) {
      const float t = l;
      l = h;
      h = t;
    }
    b[d] = l;
    b[3 + d] = h;
  }
}

// get_mcodes for float (build_radix_tree.hpp:146-176): centroid as above,
// (c - min) * inv in float, then morton32_encode (:48-65), whose scaling by
// 2^(32/Dims) and clamps are exact.
__device__ __forceinline__ void mcode_coords_f32(const double b[6], const double minc[3],
                                                 const double inv[3], double u[3]) {
  for (int d = 0; d < 3; ++d) {
    const float c = (float)(0.5 * (double)((float)b[d] + (float)b[3 + d]));
    u[d] = (double)((c - (float)minc[d]) * (float)inv[d]);
  }
}

// intersect_ray_bbox_test / intersect_ray with T = float
// (intersect_ray_impl.hpp:150-187,321-351): every step in float, tmin from
// FLT_MIN, the axis skipped when |n| <= eps.
__device__ __forceinline__ bool ray_slab_f32(float x0, float n, float bmin, float bmax,
                                             float& tmin, float& tmax, float eps) {
  if (fabsf(n) <= eps) return !((x0 < bmin) || (x0 > bmax));
  const float invn = 1.0f / n;
  float t1 = (bmin - x0) * invn;
  float t2 = (bmax - x0) * invn;
  if (t1 > t2) {
    const float t = t1;
    t1 = t2;
    t2 = t;
  }
  tmin = (t1 < tmin) ? tmin : t1;
  tmax = (t2 < tmax) ? t2 : tmax;
  return !(tmin > tmax);
}

// r: the ray {origin, unit direction} and b the box, float values held in
// doubles; dims 2 tests x and y only.
__device__ __forceinline__ bool ray_box_f32(const double r[6], const double b[6], float eps,
                                            int dims) {
  float tmin = 1.17549435e-38f;  // FLT_MIN
  float tmax = kFltMax;
  bool hit = ray_slab_f32((float)r[0], (float)r[3], (float)b[0], (float)b[3], tmin, tmax, eps);
  hit = hit && ray_slab_f32((float)r[1], (float)r[4], (float)b[1], (float)b[4], tmin, tmax, eps);
  if (dims == 3)
    hit = hit && ray_slab_f32((float)r[2], (float)r[5], (float)b[2], (float)b[5], tmin, tmax, eps);
  return hit;
}

}  // namespace axsq
