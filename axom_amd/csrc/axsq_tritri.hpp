// axsq_tritri.hpp -- triangle-triangle intersection on the device, with the
// same double operations and decisions as primal::intersect(Triangle3,
// Triangle3, includeBoundary, EPS) (primal/operators/intersect.hpp:65-71 ->
// primal/operators/detail/intersect_impl.hpp:155-382,456-471,1102-1484;
// Devillers & Guigue 2002). Corners are carried as values, never as pointers
// into a local array, so the whole test stays in registers.
#pragma once

#include "axsq_device.hpp"

namespace axsq {

struct V3 {
  double x, y, z;
};
struct V2 {
  double x, y;
};

__device__ __forceinline__ V3 v3_sub(const V3& a, const V3& b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
__device__ __forceinline__ double v3_dot(const V3& a, const V3& b) {
  return dot3(a.x, a.y, a.z, b.x, b.y, b.z);
}
__device__ __forceinline__ V3 v3_cross(const V3& u, const V3& v) {
  const double uu[3] = {u.x, u.y, u.z}, vv[3] = {v.x, v.y, v.z};
  double o[3];
  cross3(uu, vv, o);
  return {o[0], o[1], o[2]};
}
__device__ __forceinline__ V3 v3_select(int k, const V3& a, const V3& b, const V3& c) {
  return (k == 0) ? a : (k == 1) ? b : c;
}

// twoDcross, intersect_impl.hpp:538-544
__device__ __forceinline__ double tt_cross2(const V2& a, const V2& b, const V2& c) {
  return ((a.x - c.x) * (b.y - c.y) - (a.y - c.y) * (b.x - c.x));
}
// isGpeq / isLpeq against 0, fuzzy_comparators.hpp:55-84
__device__ __forceinline__ bool tt_gpeq(double v, bool inc, double eps) {
  return (inc && near_eq(v, 0.0, eps)) || is_gt(v, 0.0, eps);
}
__device__ __forceinline__ bool tt_lpeq(double v, bool inc, double eps) {
  return (inc && near_eq(v, 0.0, eps)) || is_lt(v, 0.0, eps);
}

// checkEdge, intersect_impl.hpp:1344-1392
__device__ __forceinline__ bool tt_check_edge(const V2& p1, const V2& q1, const V2& r1,
                                              const V2& p2, const V2& r2, bool inc, double eps) {
  if (tt_gpeq(tt_cross2(r2, p2, q1), inc, eps)) {
    if (!tt_gpeq(tt_cross2(r2, p1, q1), inc, eps)) return false;
    if (tt_gpeq(tt_cross2(p1, p2, q1), inc, eps)) return true;
    return tt_gpeq(tt_cross2(p1, p2, r1), inc, eps) && tt_gpeq(tt_cross2(q1, r1, p2), inc, eps);
  }
  return tt_gpeq(tt_cross2(r2, p2, r1), inc, eps) && tt_gpeq(tt_cross2(q1, r1, r2), inc, eps) &&
         tt_gpeq(tt_cross2(p1, p2, r1), inc, eps);
}

// checkVertex, intersect_impl.hpp:1394-1484
__device__ __forceinline__ bool tt_check_vertex(const V2& p1, const V2& q1, const V2& r1,
                                                const V2& p2, const V2& q2, const V2& r2,
                                                bool inc, double eps) {
  if (tt_gpeq(tt_cross2(r2, p2, q1), inc, eps)) {
    if (tt_gpeq(tt_cross2(q2, r2, q1), inc, eps)) {
      if (tt_gpeq(tt_cross2(p1, p2, q1), inc, eps)) return tt_lpeq(tt_cross2(p1, q2, q1), inc, eps);
      return tt_gpeq(tt_cross2(p1, p2, r1), inc, eps) && tt_lpeq(tt_cross2(q1, p2, r1), inc, eps);
    }
    return tt_lpeq(tt_cross2(p1, q2, q1), inc, eps) && tt_gpeq(tt_cross2(q2, r2, r1), inc, eps) &&
           tt_gpeq(tt_cross2(q1, r1, q2), inc, eps);
  }
  if (!tt_gpeq(tt_cross2(r2, p2, r1), inc, eps)) return false;
  if (tt_gpeq(tt_cross2(q1, r1, r2), inc, eps)) return tt_gpeq(tt_cross2(r1, p1, p2), inc, eps);
  return tt_gpeq(tt_cross2(q1, r1, q2), inc, eps) && tt_gpeq(tt_cross2(q2, r2, r1), inc, eps);
}

// intersectPermuted2DTriangles, intersect_impl.hpp:1281-1342: where p1 lies
// against triangle 2's three edges picks the edge or the vertex test.
__device__ __forceinline__ bool tt_permuted_2d(const V2& p1, const V2& q1, const V2& r1,
                                               const V2& p2, const V2& q2, const V2& r2,
                                               bool inc, double eps) {
  const bool a = tt_gpeq(tt_cross2(p2, q2, p1), inc, eps);
  const bool b = tt_gpeq(tt_cross2(q2, r2, p1), inc, eps);
  const bool c = tt_gpeq(tt_cross2(r2, p2, p1), inc, eps);
  if (a && b) return c ? true : tt_check_edge(p1, q1, r1, p2, r2, inc, eps);
  if (a)
    return c ? tt_check_edge(p1, q1, r1, r2, q2, inc, eps)
             : tt_check_vertex(p1, q1, r1, p2, q2, r2, inc, eps);
  if (b)
    return c ? tt_check_edge(p1, q1, r1, q2, p2, inc, eps)
             : tt_check_vertex(p1, q1, r1, q2, r2, p2, inc, eps);
  return tt_check_vertex(p1, q1, r1, r2, p2, q2, inc, eps);
}

// TriangleIntersection2D, intersect_impl.hpp:1251-1279: each triangle made
// counter-clockwise by swapping its last two corners.
__device__ __forceinline__ bool tt_tri_2d(const V2& a0, const V2& a1, const V2& a2, const V2& b0,
                                          const V2& b1, const V2& b2, bool inc, double eps) {
  const bool fa = is_lt(tt_cross2(a0, a1, a2), 0.0, eps);
  const bool fb = is_lt(tt_cross2(b0, b1, b2), 0.0, eps);
  return tt_permuted_2d(a0, fa ? a2 : a1, fa ? a1 : a2, b0, fb ? b2 : b1, fb ? b1 : b2, inc, eps);
}

// intersectCoplanar3DTriangles, intersect_impl.hpp:1190-1249: projection on
// the coordinate plane where triangle 1's normal is largest -- (z, y) with
// corners q, p, r; (x, z) with q, p, r; else (x, y) with p, q, r.
__device__ __forceinline__ bool tt_coplanar(const V3& p1, const V3& q1, const V3& r1, const V3& p2,
                                            const V3& q2, const V3& r2, const V3& n, bool inc,
                                            double eps) {
  const double n0 = fabs(n.x), n1 = fabs(n.y), n2 = fabs(n.z);
  if (is_gt(n0, n2, eps) && is_geq(n0, n1, eps)) {
    return tt_tri_2d({q1.z, q1.y}, {p1.z, p1.y}, {r1.z, r1.y}, {q2.z, q2.y}, {p2.z, p2.y},
                     {r2.z, r2.y}, inc, eps);
  }
  if (is_gt(n1, n2, eps) && is_geq(n1, n0, eps)) {
    return tt_tri_2d({q1.x, q1.z}, {p1.x, p1.z}, {r1.x, r1.z}, {q2.x, q2.z}, {p2.x, p2.z},
                     {r2.x, r2.z}, inc, eps);
  }
  return tt_tri_2d({p1.x, p1.y}, {q1.x, q1.y}, {r1.x, r1.y}, {p2.x, p2.y}, {q2.x, q2.y},
                   {r2.x, r2.y}, inc, eps);
}

// intersectTwoPermutedTriangles, intersect_impl.hpp:456-471: the interval
// test of Devillers & Guigue. Vector3(a, b) = b - a; Triangle3(a, b,
// c).normal() = (b - a) x (c - a).
__device__ __forceinline__ bool tt_two_permuted(const V3& p1, const V3& q1, const V3& r1,
                                                const V3& p2, const V3& q2, const V3& r2,
                                                bool inc, double eps) {
  const V3 n1 = v3_cross(v3_sub(p2, q1), v3_sub(p1, q1));
  if (!tt_lpeq(v3_dot(v3_sub(q2, q1), n1), inc, eps)) return false;
  const V3 n2 = v3_cross(v3_sub(p2, p1), v3_sub(r1, p1));
  return tt_lpeq(v3_dot(v3_sub(r2, p1), n2), inc, eps);
}

// The permutation step, one table for triangle 1 (intersect_impl.hpp:213-381)
// and triangle 2 (intersectOnePermutedTriangle, :1116-1187): from the signed
// distances (p, q, r) of one triangle's corners to the other's plane, the
// rotation that brings the lone corner first and whether the OTHER
// triangle's last two corners swap. False when all three are near zero.
__device__ __forceinline__ bool tt_orient(double p, double q, double r, double eps, int& rot,
                                          bool& swap) {
  if (is_gt(p, 0.0, eps)) {
    if (is_gt(q, 0.0, eps)) { rot = 2; swap = true; }
    else if (is_gt(r, 0.0, eps)) { rot = 1; swap = true; }
    else { rot = 0; swap = false; }
  } else if (is_lt(p, 0.0, eps)) {
    if (is_lt(q, 0.0, eps)) { rot = 2; swap = false; }
    else if (is_lt(r, 0.0, eps)) { rot = 1; swap = false; }
    else { rot = 0; swap = true; }
  } else if (is_lt(q, 0.0, eps)) {
    if (is_geq(r, 0.0, eps)) { rot = 1; swap = true; }
    else { rot = 0; swap = false; }
  } else if (is_gt(q, 0.0, eps)) {
    if (is_gt(r, 0.0, eps)) { rot = 0; swap = true; }
    else { rot = 1; swap = false; }
  } else if (is_gt(r, 0.0, eps)) { rot = 2; swap = false; }
  else if (is_lt(r, 0.0, eps)) { rot = 2; swap = true; }
  else return false;
  return true;
}

// nonzeroSignMatch, intersect_impl.hpp:547-554
__device__ __forceinline__ bool tt_one_side(double a, double b, double c, double eps) {
  auto sgn = [](double v) { return (0 < v) - (v < 0); };
  return !near_eq(a, 0., eps) && !near_eq(b, 0., eps) && !near_eq(c, 0., eps) && sgn(a) == sgn(b) &&
         sgn(a) == sgn(c);
}
// twoZeros || oneZeroOthersMatch, intersect_impl.hpp:556-592
__device__ __forceinline__ bool tt_touch_only(double a, double b, double c, double eps) {
  const bool za = near_eq(a, 0.0, eps), zb = near_eq(b, 0.0, eps), zc = near_eq(c, 0.0, eps);
  const int nz = (int)za + (int)zb + (int)zc;
  if (nz == 2) return true;
  return nz == 1 && ((za && is_gt(b * c, 0.0, eps)) || (zb && is_gt(c * a, 0.0, eps)) ||
                     (zc && is_gt(a * b, 0.0, eps)));
}

// Triangle::normal().unitVector() of (a, b, c) (Triangle.hpp:99-102,
// Vector.hpp:477-490)
__device__ __forceinline__ V3 tt_unit_normal(const V3& a, const V3& b, const V3& c) {
  const V3 n = v3_cross(v3_sub(b, a), v3_sub(c, a));
  const double nn[3] = {n.x, n.y, n.z};
  double u[3];
  unit3(nn, u);
  return {u[0], u[1], u[2]};
}

// primal::intersect(Triangle3 (a0, a1, a2), Triangle3 (b0, b1, b2), inc, eps)
__device__ __forceinline__ bool tri_tri_intersect(const V3& a0, const V3& a1, const V3& a2,
                                                  const V3& b0, const V3& b1, const V3& b2,
                                                  bool inc, double eps) {
  const V3 nb = tt_unit_normal(b0, b1, b2);
  const double da0 = v3_dot(v3_sub(a0, b2), nb), da1 = v3_dot(v3_sub(a1, b2), nb),
               da2 = v3_dot(v3_sub(a2, b2), nb);
  if (tt_one_side(da0, da1, da2, eps)) return false;
  if (!inc && tt_touch_only(da0, da1, da2, eps)) return false;
  const V3 na = tt_unit_normal(a0, a1, a2);
  const double db0 = v3_dot(v3_sub(b0, a2), na), db1 = v3_dot(v3_sub(b1, a2), na),
               db2 = v3_dot(v3_sub(b2, a2), na);
  if (tt_one_side(db0, db1, db2, eps)) return false;
  if (!inc && tt_touch_only(db0, db1, db2, eps)) return false;
  int rot_a;
  bool swap_b;
  if (!tt_orient(da0, da1, da2, eps, rot_a, swap_b))
    return tt_coplanar(a0, a1, a2, b0, b1, b2, na, inc, eps);
  const V3 P1 = v3_select(rot_a, a0, a1, a2);
  const V3 Q1 = v3_select(rot_a, a1, a2, a0);
  const V3 R1 = v3_select(rot_a, a2, a0, a1);
  const V3 B1 = swap_b ? b2 : b1, B2 = swap_b ? b1 : b2;
  const double e1 = swap_b ? db2 : db1, e2 = swap_b ? db1 : db2;
  int rot_b;
  bool swap_a;
  if (!tt_orient(db0, e1, e2, eps, rot_b, swap_a))
    return tt_coplanar(P1, Q1, R1, b0, B1, B2, na, inc, eps);
  return tt_two_permuted(P1, swap_a ? R1 : Q1, swap_a ? Q1 : R1, v3_select(rot_b, b0, B1, B2),
                         v3_select(rot_b, B1, B2, b0), v3_select(rot_b, B2, b0, B1), inc, eps);
}

}  // namespace axsq
