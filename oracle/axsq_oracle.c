/*
 * axsq_oracle.c -- CPU restatement of Axom's BVH + SignedDistance SEQ path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the parity checker and the timed CPU
 * baseline; the HIP product path never calls it. Every function cites the
 * reference file:line it follows (paths relative to the reference's
 * src/axom/). Compile with -ffp-contract=off and without fast-math.
 *
 * Parity pins (what checks each function): see tests/test_oracle_*.py --
 * reference known answers (spin_bvh.cpp, spin_morton.cpp,
 * primal_closest_point.cpp, quest_signed_distance*.cpp), brute force,
 * closed forms and tree invariants.
 */
#include "axsq_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define ORC_BARRIER (-2000000000) /* bvh_traverse.hpp:80 */
#define ORC_STACK 64              /* bvh_traverse.hpp:79 */

/* ------------------------------------------------------------------------ */
/* primal subset                                                            */
/* ------------------------------------------------------------------------ */

/* core/utilities/Utilities.hpp:43-46 (abs) and :318-321 (isNearlyEqual). */
static inline int near_eq(double a, double b, double thresh)
{
  double d = a - b;
  d = (d < 0.0) ? -d : d;
  return d <= thresh;
}

/* primal/operators/detail/fuzzy_comparators.hpp:26-44 */
static inline int is_gt(double x, double y, double eps) { return (x > y) && !near_eq(x, y, eps); }
static inline int is_lt(double x, double y, double eps) { return (x < y) && !near_eq(x, y, eps); }
static inline int is_leq(double x, double y, double eps) { return !is_gt(x, y, eps); }
static inline int is_geq(double x, double y, double eps) { return !is_lt(x, y, eps); }

/* primal/geometry/Vector.hpp:543-552: ((0 + u0 v0) + u1 v1) + u2 v2 */
static inline double dot3(const double u[3], const double v[3])
{
  double r = 0.0;
  r += u[0] * v[0];
  r += u[1] * v[1];
  r += u[2] * v[2];
  return r;
}

/* primal/geometry/Vector.hpp:564-571 via core/numerics/Determinants.hpp:31-38 */
static inline void cross3(const double u[3], const double v[3], double out[3])
{
  out[0] = u[1] * v[2] - v[1] * u[2];
  out[1] = v[0] * u[2] - u[0] * v[2];
  out[2] = u[0] * v[1] - v[0] * u[1];
}

/* primal/geometry/Vector.hpp:477-490; division is multiplication by the
 * reciprocal (core/NumericArray.hpp:510-514); PRIMAL_TINY = 1e-50. */
void orc_unit_vector(const double v[3], double out[3])
{
  const double len_sq = dot3(v, v);
  if(len_sq >= 1e-50)
  {
    const double inv = 1. / sqrt(len_sq);
    out[0] = v[0] * inv;
    out[1] = v[1] * inv;
    out[2] = v[2] * inv;
  }
  else
  {
    out[0] = 1.0;
    out[1] = 0.0;
    out[2] = 0.0;
  }
}

/* primal/geometry/Ray.hpp:122-128: the Ray stores direction.unitVector(). */
void orc_ray_normalize(const double d[3], double out[3]) { orc_unit_vector(d, out); }

/* primal/geometry/BoundingBox.hpp:451-462. The box operations take the
 * dimension nd (2 or 3) of BoundingBox<double, NDIMS>; a 2-D box is held in
 * orc_box with its z entries unused. */
static inline int box_valid_d(const orc_box* b, int nd)
{
  for(int d = 0; d < nd; ++d)
    if(b->min[d] > b->max[d]) return 0;
  return 1;
}
int orc_box_valid(const orc_box* b) { return box_valid_d(b, 3); }

/* BoundingBox.hpp:72-73,588-592 */
static inline void box_clear(orc_box* b)
{
  for(int d = 0; d < 3; ++d)
  {
    b->min[d] = DBL_MAX;
    b->max[d] = -DBL_MAX;
  }
}

/* BoundingBox.hpp:466-484 */
static inline void box_add_point_d(orc_box* b, const double p[3], int nd)
{
  for(int d = 0; d < nd; ++d)
  {
    if(p[d] < b->min[d]) b->min[d] = p[d];
    if(p[d] > b->max[d]) b->max[d] = p[d];
  }
}

/* BoundingBox.hpp:486-508: an invalid receiver is overwritten, an invalid
 * argument is ignored by a valid receiver. */
static inline void box_add_box_d(orc_box* b, const orc_box* o, int nd)
{
  if(box_valid_d(b, nd))
  {
    if(box_valid_d(o, nd))
    {
      box_add_point_d(b, o->min, nd);
      box_add_point_d(b, o->max, nd);
    }
  }
  else
  {
    *b = *o;
  }
}
void orc_box_add_box(orc_box* b, const orc_box* o) { box_add_box_d(b, o, 3); }

/* BoundingBox.hpp:547-561 with Point::midpoint (Point.hpp:279-290) and
 * range() = max - min; fixBounds at :575-585. */
static inline void box_scale_d(orc_box* b, double sf, int nd)
{
  if(!box_valid_d(b, nd)) return;
  const double s = sf * 0.5;
  double mid[3], r[3];
  for(int d = 0; d < nd; ++d)
  {
    mid[d] = 0.5 * (b->min[d] + b->max[d]);
    r[d] = (b->max[d] - b->min[d]) * s;
  }
  for(int d = 0; d < nd; ++d)
  {
    b->min[d] = mid[d] - r[d];
    b->max[d] = mid[d] + r[d];
    if(b->min[d] > b->max[d])
    {
      const double t = b->min[d];
      b->min[d] = b->max[d];
      b->max[d] = t;
    }
  }
}
void orc_box_scale(orc_box* b, double sf) { box_scale_d(b, sf, 3); }

/* BoundingBox.hpp:389-401 */
static inline int box_contains_d(const orc_box* b, const double p[3], int nd)
{
  for(int d = 0; d < nd; ++d)
    if(p[d] < b->min[d] || p[d] > b->max[d]) return 0;
  return 1;
}
static inline int box_contains(const orc_box* b, const double p[3]) { return box_contains_d(b, p, 3); }

/* BoundingBox.hpp:433-447 + intersect_bounding_box_impl.hpp:34-41;
 * q is the query box (this), b the tree box (otherBB). */
static inline int box_intersects_d(const orc_box* q, const orc_box* b, int nd)
{
  for(int d = 0; d < nd; ++d)
    if(!(q->max[d] >= b->min[d] && q->min[d] <= b->max[d])) return 0;
  return 1;
}

/* primal/operators/squared_distance.hpp:77-102 (Point, Box) and :62-66. */
double orc_sqdist_point_box(const double p[3], const orc_box* b)
{
  if(!orc_box_valid(b)) return DBL_MAX;
  if(box_contains(b, p)) return 0;
  double v[3];
  for(int d = 0; d < 3; ++d)
  {
    /* core/utilities/Utilities.hpp:147-150 clampVal */
    const double c = (p[d] < b->min[d]) ? b->min[d] : (p[d] > b->max[d]) ? b->max[d] : p[d];
    v[d] = c - p[d];
  }
  return dot3(v, v);
}

static inline double sqdist_pp(const double a[3], const double b[3])
{
  double v[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  return dot3(v, v);
}

/* primal/operators/detail/intersect_ray_impl.hpp:150-187 (per-axis slab),
 * utilities max/min at core/utilities/Utilities.hpp:80-96. */
static inline int ray_slab(double x0, double n, double bmin, double bmax, double* tmin,
                           double* tmax, double eps)
{
  if(near_eq(n, 0.0, eps))
  {
    return ((x0 < bmin) || (x0 > bmax)) ? 0 : 1;
  }
  const double invn = 1.0 / n;
  double t1 = (bmin - x0) * invn;
  double t2 = (bmax - x0) * invn;
  if(t1 > t2)
  {
    const double t = t1;
    t1 = t2;
    t2 = t;
  }
  *tmin = (t1 < *tmin) ? *tmin : t1;
  *tmax = (t2 < *tmax) ? t2 : *tmax;
  return (*tmin > *tmax) ? 0 : 1;
}

/* intersect_ray_impl.hpp:321-351: tmin starts at DBL_MIN (a ray, not a line),
 * the && short-circuits the remaining axes. */
static inline int ray_box_d(const double o[3], const double dir[3], const orc_box* b, double eps,
                            int nd)
{
  double tmin = DBL_MIN, tmax = DBL_MAX;
  int hit = 1;
  for(int d = 0; d < nd; ++d)
    hit = hit && ray_slab(o[d], dir[d], b->min[d], b->max[d], &tmin, &tmax, eps);
  return hit;
}
int orc_ray_box(const double o[3], const double dir[3], const orc_box* b, double eps)
{
  return ray_box_d(o, dir, b, eps, 3);
}

/* ---- FloatType = float (spin::BVH<NDIMS, Exec, float>) -------------------
 * A float tree's boxes are float values held in doubles (exact); unions,
 * comparisons and the point and box predicates then give the float results
 * as they are. These steps round to float where BoundingBox<float> and the
 * float instantiation of the build and of intersect_ray do. */

/* BoundingBox<float>::scale (BoundingBox.hpp:547-561): midpoint
 * float(0.5 * (min + max)) with the sum in float (Point.hpp:279-290), radius
 * float(range * float(sf * 0.5)) from NumericArray::operator*=(double)
 * (core/NumericArray.hpp:498-506; a product of floats is exact in double),
 * then min/max in float and fixBounds. */
static inline void box_scale_f(orc_box* b, float sf, int nd)
{
  if(!box_valid_d(b, nd)) return;
  const float s = (float)((double)sf * 0.5);
  float lo[3], hi[3];
  for(int d = 0; d < nd; ++d)
  {
    const float mn = (float)b->min[d];
    const float mx = (float)b->max[d];
    const float sum = mn + mx;
    const float mid = (float)(0.5 * (double)sum);
    const float range = mx - mn;
    const float r = (float)((double)range * (double)s);
    lo[d] = mid - r;
    hi[d] = mid + r;
  }
  for(int d = 0; d < nd; ++d)
  {
    if(lo[d] > hi[d])
    {
      const float t = lo[d];
      lo[d] = hi[d];
      hi[d] = t;
    }
    b->min[d] = lo[d];
    b->max[d] = hi[d];
  }
}

/* get_mcodes + morton32_encode with FloatType = float (build_radix_tree.hpp:
 * 48-65,146-176): (centroid - min) * inv_extent in float. */
static inline uint32_t box_mcode_f(const orc_box* b, const float minc[3], const float inv[3],
                                   int nd)
{
  const float scale = (nd == 2) ? 65536.0f : 1024.0f;
  int32_t q[3] = {0, 0, 0};
  for(int d = 0; d < nd; ++d)
  {
    const float sum = (float)b->min[d] + (float)b->max[d];
    const float c = (float)(0.5 * (double)sum);
    const float u = (c - minc[d]) * inv[d];
    q[d] = (int32_t)fminf(fmaxf(u * scale, 0.0f), scale - 1.0f);
  }
  return (nd == 2) ? orc_morton2(q[0], q[1]) : orc_morton3(q[0], q[1], q[2]);
}

/* intersect_ray_bbox_test / intersect_ray with T = float (intersect_ray_impl.hpp:
 * 150-187,321-351): isNearlyEqual(n, 0, EPS) as |n - 0| <= EPS, 1 / n and the
 * slab parameters in float, tmin from FLT_MIN. */
static inline int ray_slab_f(float x0, float n, float bmin, float bmax, float* tmin, float* tmax,
                             float eps)
{
  if(fabsf(n - 0.0f) <= eps)
  {
    return ((x0 < bmin) || (x0 > bmax)) ? 0 : 1;
  }
  const float invn = 1.0f / n;
  float t1 = (bmin - x0) * invn;
  float t2 = (bmax - x0) * invn;
  if(t1 > t2)
  {
    const float t = t1;
    t1 = t2;
    t2 = t;
  }
  *tmin = (t1 < *tmin) ? *tmin : t1;
  *tmax = (t2 < *tmax) ? t2 : *tmax;
  return (*tmin > *tmax) ? 0 : 1;
}

static inline int ray_box_f(const double o[3], const double dir[3], const orc_box* b, float eps,
                            int nd)
{
  float tmin = FLT_MIN, tmax = FLT_MAX;
  int hit = 1;
  for(int d = 0; d < nd; ++d)
    hit = hit &&
      ray_slab_f((float)o[d], (float)dir[d], (float)b->min[d], (float)b->max[d], &tmin, &tmax,
                 eps);
  return hit;
}
int orc_ray_box_f32(const double o[3], const double dir[3], const orc_box* b, float eps)
{
  return ray_box_f(o, dir, b, eps, 3);
}

/* primal/operators/closest_point.hpp:162-290 (Ericson RTCD 5.1.5) with the
 * degenerate-edge guard at :215. Returns the location code:
 * 0,1,2 vertex A,B,C; -1 edge AB; -3 edge AC; -2 edge BC; 3 face. */
int orc_closest_point(const double P[3], const double A[3], const double B[3],
                      const double C[3], double EPS, double cp[3])
{
  double ab[3], ac[3], ap[3], bp[3], cpv[3];
  for(int d = 0; d < 3; ++d)
  {
    ab[d] = B[d] - A[d];
    ac[d] = C[d] - A[d];
    ap[d] = P[d] - A[d];
  }
  const double d1 = dot3(ab, ap);
  const double d2 = dot3(ac, ap);
  if(is_leq(d1, 0.0, EPS) && is_leq(d2, 0.0, EPS))
  {
    memcpy(cp, A, 3 * sizeof(double));
    return 0;
  }
  for(int d = 0; d < 3; ++d) bp[d] = P[d] - B[d];
  const double d3 = dot3(ab, bp);
  const double d4 = dot3(ac, bp);
  if(is_geq(d3, 0.0, EPS) && is_leq(d4, d3, EPS))
  {
    memcpy(cp, B, 3 * sizeof(double));
    return 1;
  }
  const double vc = d1 * d4 - d3 * d2;
  if(is_leq(vc, 0.0, EPS) && is_geq(d1, 0.0, EPS) && is_leq(d3, 0.0, EPS) &&
     !near_eq(d1, d3, EPS))
  {
    const double v = d1 / (d1 - d3);
    for(int d = 0; d < 3; ++d) cp[d] = A[d] + ab[d] * v;
    return -1;
  }
  for(int d = 0; d < 3; ++d) cpv[d] = P[d] - C[d];
  const double d5 = dot3(ab, cpv);
  const double d6 = dot3(ac, cpv);
  if(is_geq(d6, 0.0, EPS) && is_leq(d5, d6, EPS))
  {
    memcpy(cp, C, 3 * sizeof(double));
    return 2;
  }
  const double vb = d5 * d2 - d1 * d6;
  if(is_leq(vb, 0.0, EPS) && is_geq(d2, 0.0, EPS) && is_leq(d6, 0.0, EPS))
  {
    const double w = d2 / (d2 - d6);
    for(int d = 0; d < 3; ++d) cp[d] = A[d] + ac[d] * w;
    return -3;
  }
  const double va = d3 * d6 - d5 * d4;
  if(is_leq(va, 0.0, EPS) && is_geq(d4 - d3, 0.0, EPS) && is_geq(d5 - d6, 0.0, EPS))
  {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    for(int d = 0; d < 3; ++d) cp[d] = B[d] + (C[d] - B[d]) * w;
    return -2;
  }
  const double denom = 1.0 / (va + vb + vc);
  const double v = vb * denom;
  const double w = vc * denom;
  for(int d = 0; d < 3; ++d)
  {
    const double n = ab[d] * v + ac[d] * w;
    cp[d] = A[d] + n;
  }
  return 3;
}

/* ------------------------------------------------------------------------ */
/* BVH build: spin/internal/linear_bvh/build_radix_tree.hpp                 */
/* ------------------------------------------------------------------------ */

/* 10-bit spread of spin/MortonIndex.hpp:151-159,369-470 restated with the
 * classic shift/mask sequence for 3-D, 10 bits per axis. */
static inline uint32_t spread10(uint32_t v)
{
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

uint32_t orc_morton3(int32_t x, int32_t y, int32_t z)
{
  return spread10((uint32_t)x) | (spread10((uint32_t)y) << 1) | (spread10((uint32_t)z) << 2);
}

/* The 2-D Mortonizer (spin/MortonIndex.hpp:226-318): expandBits of an int32
 * runs the shift/mask pairs 16/0x0000ffff, 8/0x00ff00ff, 4/0x0f0f0f0f,
 * 2/0x33333333, 1/0x55555555; x in the even bits, y in the odd ones. */
static inline uint32_t spread16(uint32_t v)
{
  v = (v | (v << 16)) & 0x0000ffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

uint32_t orc_morton2(int32_t x, int32_t y) { return spread16((uint32_t)x) | (spread16((uint32_t)y) << 1); }

/* build_radix_tree.hpp:48-65 (morton32_encode: 32 / Dims bits per axis, so
 * 10 in 3-D and 16 in 2-D) on the normalised centroid computed at :154-175. */
static inline uint32_t box_mcode(const orc_box* b, const double minc[3], const double inv[3], int nd)
{
  const double scale = (nd == 2) ? 65536.0 : 1024.0;
  int32_t q[3];
  for(int d = 0; d < nd; ++d)
  {
    const double c = 0.5 * (b->min[d] + b->max[d]);
    const double u = (c - minc[d]) * inv[d];
    q[d] = (int32_t)fmin(fmax(u * scale, 0.0), scale - 1.0);
  }
  return (nd == 2) ? orc_morton2(q[0], q[1]) : orc_morton3(q[0], q[1], q[2]);
}

/* build_radix_tree.hpp:265-287 */
static inline int32_t karras_delta(int32_t a, int32_t b, int32_t inner, const uint32_t* mc)
{
  const int oor = (b < 0 || b > inner);
  const int32_t bb = oor ? 0 : b;
  uint32_t ex = mc[a] ^ mc[bb];
  const int tie = (ex == 0);
  if(tie) ex = (uint32_t)a ^ (uint32_t)bb;
  int32_t cnt = ex ? __builtin_clz(ex) : 32;
  if(tie) cnt += 32;
  return oor ? -1 : cnt;
}

/* Stable LSD radix sort of (key, idx); equals std::stable_sort by key
 * (build_radix_tree.hpp:243-260) because both give the unique
 * (key, original index) lexicographic order. */
static void stable_sort_codes(uint32_t* keys, int32_t* idx, int32_t n)
{
  uint32_t* k2 = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
  int32_t* i2 = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  for(int shift = 0; shift < 32; shift += 8)
  {
    size_t cnt[257];
    memset(cnt, 0, sizeof(cnt));
    for(int32_t i = 0; i < n; ++i) cnt[((keys[i] >> shift) & 0xffu) + 1]++;
    for(int b = 0; b < 256; ++b) cnt[b + 1] += cnt[b];
    for(int32_t i = 0; i < n; ++i)
    {
      const size_t dst = cnt[(keys[i] >> shift) & 0xffu]++;
      k2[dst] = keys[i];
      i2[dst] = idx[i];
    }
    memcpy(keys, k2, sizeof(uint32_t) * (size_t)n);
    memcpy(idx, i2, sizeof(int32_t) * (size_t)n);
  }
  free(k2);
  free(i2);
}

/* boxes_in: AoS of 2 * nd doubles (min, max) per box; f32: a float tree over
 * boxes_f (2 * nd floats per box) with a float scale factor. */
static orc_bvh* bvh_new_any(const double* boxes_in, const float* boxes_f, int32_t n_in,
                            double scale, int nd, int f32)
{
  orc_bvh* t = (orc_bvh*)calloc(1, sizeof(orc_bvh));
  t->dims = nd;
  t->f32 = f32;
  /* spin/BVH.hpp:441-464: pad 0 or 1 boxes to 2 with an invalid box. */
  const int32_t n = (n_in <= 1) ? 2 : n_in;
  const int32_t inner = n - 1;
  t->n = n;
  t->mcodes = (uint32_t*)malloc(sizeof(uint32_t) * n);
  t->leafs = (int32_t*)malloc(sizeof(int32_t) * n);
  t->lchild = (int32_t*)malloc(sizeof(int32_t) * inner);
  t->rchild = (int32_t*)malloc(sizeof(int32_t) * inner);
  t->parents = (int32_t*)malloc(sizeof(int32_t) * (2 * n - 1));
  t->inner_aabbs = (orc_box*)malloc(sizeof(orc_box) * inner);
  t->leaf_aabbs = (orc_box*)malloc(sizeof(orc_box) * n);
  t->nodes = (orc_box*)malloc(sizeof(orc_box) * 2 * inner);
  t->children = (int32_t*)malloc(sizeof(int32_t) * 2 * inner);

  /* transform_boxes :85-99 */
  orc_box* aabb = (orc_box*)malloc(sizeof(orc_box) * n);
  for(int32_t i = 0; i < n; ++i)
  {
    box_clear(&aabb[i]);
    if(f32)
    {
      /* BoundingBox<float>'s invalid box (BoundingBox.hpp:72-73) */
      for(int d = 0; d < 3; ++d)
      {
        aabb[i].min[d] = FLT_MAX;
        aabb[i].max[d] = -FLT_MAX;
      }
    }
    if(n_in > 1 || (n_in == 1 && i == 0))
    {
      for(int d = 0; d < 3; ++d)
      {
        if(f32)
        {
          const float* b = boxes_f + 2 * nd * (size_t)i;
          aabb[i].min[d] = (d < nd) ? (double)b[d] : 0.0;
          aabb[i].max[d] = (d < nd) ? (double)b[nd + d] : 0.0;
        }
        else
        {
          const double* b = boxes_in + 2 * nd * (size_t)i;
          aabb[i].min[d] = (d < nd) ? b[d] : 0.0;
          aabb[i].max[d] = (d < nd) ? b[nd + d] : 0.0;
        }
      }
    }
    if(f32)
      box_scale_f(&aabb[i], (float)scale, nd);
    else
      box_scale_d(&aabb[i], scale, nd);
  }

  /* reduce :134-142 (SEQ path: addBox from an invalid box) */
  box_clear(&t->bounds);
  for(int32_t i = 0; i < n; ++i) box_add_box_d(&t->bounds, &aabb[i], nd);

  /* get_mcodes :146-176; 1.f/extent promotes to double, thresh 1e-8. */
  double minc[3], inv[3];
  for(int d = 0; d < 3; ++d)
  {
    const double ext = t->bounds.max[d] - t->bounds.min[d];
    minc[d] = t->bounds.min[d];
    inv[d] = near_eq(ext, 0.0, 1.0e-8) ? 0.0 : 1.0 / ext;
  }
  /* the float tree: extent and 1.f / extent in float, the threshold a float */
  float minc_f[3], inv_f[3];
  for(int d = 0; d < 3; ++d)
  {
    const float ext = (float)t->bounds.max[d] - (float)t->bounds.min[d];
    minc_f[d] = (float)t->bounds.min[d];
    inv_f[d] = (fabsf(ext - 0.0f) <= (float)1.0e-8) ? 0.0f : 1.0f / ext;
  }
  for(int32_t i = 0; i < n; ++i)
  {
    t->mcodes[i] = f32 ? box_mcode_f(&aabb[i], minc_f, inv_f, nd)
                       : box_mcode(&aabb[i], minc, inv, nd);
    t->leafs[i] = i;
  }

  /* sort_mcodes + reorder :243-260, :606 */
  stable_sort_codes(t->mcodes, t->leafs, n);
  for(int32_t i = 0; i < n; ++i) t->leaf_aabbs[i] = aabb[t->leafs[i]];
  free(aabb);

  /* build_tree :290-384 (Karras 2012) */
  const uint32_t* mc = t->mcodes;
  for(int32_t i = 0; i < inner; ++i)
  {
    const int32_t d =
      (karras_delta(i, i + 1, inner, mc) - karras_delta(i, i - 1, inner, mc)) < 0 ? -1 : 1;
    const int32_t dmin = karras_delta(i, i - d, inner, mc);
    int32_t lmax = 2;
    while(karras_delta(i, i + lmax * d, inner, mc) > dmin) lmax *= 2;
    int32_t l = 0;
    for(int32_t s = lmax / 2; s >= 1; s /= 2)
      if(karras_delta(i, i + (l + s) * d, inner, mc) > dmin) l += s;
    const int32_t j = i + l * d;
    const int32_t dnode = karras_delta(i, j, inner, mc);
    int32_t s = 0;
    double div = 2.0;
    int32_t step = (int32_t)ceil((double)(float)l / div);
    for(;;)
    {
      if(karras_delta(i, i + (s + step) * d, inner, mc) > dnode) s += step;
      if(step == 1) break;
      div *= 2;
      step = (int32_t)ceil((double)(float)l / div);
    }
    const int32_t split = i + s * d + (d < 0 ? d : 0);
    const int32_t lo = (i < j) ? i : j;
    const int32_t hi = (i < j) ? j : i;
    if(lo == split)
    {
      t->parents[split + inner] = i;
      t->lchild[i] = split + inner;
    }
    else
    {
      t->parents[split] = i;
      t->lchild[i] = split;
    }
    if(hi == split + 1)
    {
      t->parents[split + inner + 1] = i;
      t->rchild[i] = split + inner + 1;
    }
    else
    {
      t->parents[split + 1] = i;
      t->rchild[i] = split + 1;
    }
    if(i == 0) t->parents[0] = -1;
  }

  /* propagate_aabbs :505-576 in SEQ leaf order */
  int32_t* counters = (int32_t*)calloc((size_t)inner, sizeof(int32_t));
  for(int32_t i = 0; i < inner; ++i) box_clear(&t->inner_aabbs[i]);
  for(int32_t i = 0; i < n; ++i)
  {
    orc_box acc = t->leaf_aabbs[i];
    int32_t last = inner + i;
    int32_t cur = t->parents[inner + i];
    while(cur != -1)
    {
      const int32_t old = counters[cur]++;
      if(old == 0) break;
      const int32_t lc = t->lchild[cur], rc = t->rchild[cur];
      const int32_t other = (lc == last) ? rc : lc;
      const orc_box ob = (other >= inner) ? t->leaf_aabbs[other - inner] : t->inner_aabbs[other];
      box_add_box_d(&acc, &ob, nd);
      t->inner_aabbs[cur] = acc;
      last = cur;
      cur = t->parents[cur];
    }
  }
  free(counters);

  /* emit_bvh_parents, spin/policy/LinearBVH.hpp:225-264 */
  for(int32_t i = 0; i < inner; ++i)
  {
    int32_t lc = t->lchild[i], rc = t->rchild[i];
    orc_box lb, rb;
    if(lc >= inner)
    {
      lb = t->leaf_aabbs[lc - inner];
      lc = -(lc - inner + 1);
    }
    else
    {
      lb = t->inner_aabbs[lc];
      lc *= 2;
    }
    if(rc >= inner)
    {
      rb = t->leaf_aabbs[rc - inner];
      rc = -(rc - inner + 1);
    }
    else
    {
      rb = t->inner_aabbs[rc];
      rc *= 2;
    }
    t->nodes[2 * i] = lb;
    t->nodes[2 * i + 1] = rb;
    t->children[2 * i] = lc;
    t->children[2 * i + 1] = rc;
  }
  return t;
}

orc_bvh* orc_bvh_new(const double* boxes_in, int32_t n_in, double scale)
{
  return bvh_new_any(boxes_in, NULL, n_in, scale, 3, 0);
}

orc_bvh* orc_bvh_new2d(const double* boxes_in, int32_t n_in, double scale)
{
  return bvh_new_any(boxes_in, NULL, n_in, scale, 2, 0);
}

orc_bvh* orc_bvh_new_f32(const float* boxes_in, int32_t n_in, int nd, float scale)
{
  return bvh_new_any(NULL, boxes_in, n_in, (double)scale, nd, 1);
}

void orc_bvh_free(orc_bvh* t)
{
  if(!t) return;
  free(t->mcodes);
  free(t->leafs);
  free(t->lchild);
  free(t->rchild);
  free(t->parents);
  free(t->inner_aabbs);
  free(t->leaf_aabbs);
  free(t->nodes);
  free(t->children);
  free(t);
}

int32_t orc_bvh_size(const orc_bvh* t) { return t->n; }

void orc_bvh_export(const orc_bvh* t, double* bounds6, uint32_t* mcodes, int32_t* leafs,
                    int32_t* lchild, int32_t* rchild, int32_t* parents, double* inner_aabbs,
                    double* nodes, int32_t* children)
{
  const int32_t n = t->n, inner = n - 1;
  if(bounds6) memcpy(bounds6, &t->bounds, sizeof(orc_box));
  if(mcodes) memcpy(mcodes, t->mcodes, sizeof(uint32_t) * n);
  if(leafs) memcpy(leafs, t->leafs, sizeof(int32_t) * n);
  if(lchild) memcpy(lchild, t->lchild, sizeof(int32_t) * inner);
  if(rchild) memcpy(rchild, t->rchild, sizeof(int32_t) * inner);
  if(parents) memcpy(parents, t->parents, sizeof(int32_t) * (2 * n - 1));
  if(inner_aabbs) memcpy(inner_aabbs, t->inner_aabbs, sizeof(orc_box) * inner);
  if(nodes) memcpy(nodes, t->nodes, sizeof(orc_box) * 2 * inner);
  if(children) memcpy(children, t->children, sizeof(int32_t) * 2 * inner);
}

/* ------------------------------------------------------------------------ */
/* Traversal: spin/internal/linear_bvh/bvh_traverse.hpp:66-154               */
/* ------------------------------------------------------------------------ */

typedef int (*pred_fn)(const void* prim, const orc_box* b, void* ctx);
typedef void (*leaf_fn)(int32_t pos, const int32_t* leafs, void* ctx);
typedef int (*pref_fn)(const orc_box* l, const orc_box* r, const void* prim);

/* Returns the number of inner-node visits; -1 on stack overflow (the
 * reference has no bound check, bvh_traverse.hpp:79-120). When max_sp is not
 * NULL it is raised to the deepest stack entry the walk pushed. */
static inline int32_t traverse(const orc_bvh* t, const void* prim, pred_fn B, void* pctx,
                               leaf_fn A, void* actx, pref_fn Comp, int32_t* max_sp)
{
  int32_t todo[ORC_STACK];
  int32_t sp = 0;
  int32_t visits = 0;
  todo[0] = ORC_BARRIER;
  int32_t found = 0;
  int32_t cur = 0;
  while(cur != ORC_BARRIER)
  {
    while(cur >= 0)
    {
      const orc_box* lb = &t->nodes[cur];
      const orc_box* rb = &t->nodes[cur + 1];
      ++visits;
      const int in_l = orc_box_valid(lb) ? B(prim, lb, pctx) : 0;
      const int in_r = orc_box_valid(rb) ? B(prim, rb, pctx) : 0;
      int32_t lc = t->children[cur];
      int32_t rc = t->children[cur + 1];
      const int swap = Comp(lb, rb, prim);
      if(!in_l && !in_r)
      {
        cur = todo[sp];
        sp--;
      }
      else
      {
        cur = in_l ? lc : rc;
        if(in_l && in_r)
        {
          if(swap)
          {
            const int32_t tmp = cur;
            cur = rc;
            rc = tmp;
          }
          sp++;
          if(sp >= ORC_STACK) return -1;
          if(max_sp && sp > *max_sp) *max_sp = sp;
          todo[sp] = rc;
        }
      }
      if(cur < 0 && !(found < 0))
      {
        found = cur;
        if(cur != ORC_BARRIER)
        {
          cur = todo[sp];
          sp--;
        }
      }
    }
    while(found < 0 && found != ORC_BARRIER)
    {
      A(-found - 1, t->leafs, actx);
      found = cur;
      if(cur < 0 && cur != ORC_BARRIER)
      {
        cur = todo[sp];
        sp--;
      }
    }
    found = 0;
  }
  return visits;
}

static int no_pref(const orc_box* l, const orc_box* r, const void* p)
{
  (void)l;
  (void)r;
  (void)p;
  return 0;
}

/* LinearBVH.hpp:75-82: near-first by squared distance to the centroids; an
 * invalid right box never wins. */
static int point_pref(const orc_box* l, const orc_box* r, const void* prim)
{
  const double* p = (const double*)prim;
  double cl[3], cr[3];
  for(int d = 0; d < 3; ++d) cl[d] = 0.5 * (l->min[d] + l->max[d]);
  const double dl = sqdist_pp(p, cl);
  double dr = DBL_MAX;
  if(orc_box_valid(r))
  {
    for(int d = 0; d < 3; ++d) cr[d] = 0.5 * (r->min[d] + r->max[d]);
    dr = sqdist_pp(p, cr);
  }
  return dl > dr;
}

typedef struct
{
  double tol;
  int nd;
  int f32; /* a float tree: the ray test in float with a float tolerance */
} find_ctx;

/* prim: a point (3 doubles), a box (orc_box) or a ray (origin, direction),
 * with the unused z entries of a 2-D query set to 0 */
static int pred_point(const void* prim, const orc_box* b, void* ctx)
{
  return box_contains_d(b, (const double*)prim, ((find_ctx*)ctx)->nd);
}
static int pred_box(const void* prim, const orc_box* b, void* ctx)
{
  return box_intersects_d((const orc_box*)prim, b, ((find_ctx*)ctx)->nd);
}
static int pred_ray(const void* prim, const orc_box* b, void* ctx)
{
  const double* r = (const double*)prim;
  const find_ctx* c = (const find_ctx*)ctx;
  return c->f32 ? ray_box_f(r, r + 3, b, (float)c->tol, c->nd) : ray_box_d(r, r + 3, b, c->tol, c->nd);
}

/* query i of an AoS array of nd-dimensional points (nd doubles), boxes or
 * rays (2 * nd doubles) as the 3-D layout the predicates read */
static void load_query(const double* prims, int kind, int nd, int64_t i, double q[6])
{
  const int w = (kind == 0) ? nd : 2 * nd;
  const double* p = prims + (size_t)w * (size_t)i;
  for(int k = 0; k < 6; ++k) q[k] = 0.0;
  if(kind == 0)
  {
    for(int d = 0; d < nd; ++d) q[d] = p[d];
  }
  else
  {
    for(int d = 0; d < nd; ++d)
    {
      q[d] = p[d];
      q[3 + d] = p[nd + d];
    }
  }
}

typedef struct
{
  int32_t count;
  int32_t* out;
} collect_ctx;

static void leaf_count(int32_t pos, const int32_t* leafs, void* ctx)
{
  collect_ctx* c = (collect_ctx*)ctx;
  if(c->out) c->out[c->count] = leafs[pos];
  c->count++;
}

typedef struct
{
  const orc_bvh* t;
  int kind;
  const double* prims;
  double tol;
  int32_t* counts;
  const int32_t* offsets;
  int32_t* cands;
  int64_t begin, end;
  int overflow;
} find_job;

static void* find_worker(void* arg)
{
  find_job* j = (find_job*)arg;
  pred_fn B = (j->kind == 0) ? pred_point : (j->kind == 1) ? pred_box : pred_ray;
  find_ctx fc = {j->tol, j->t->dims, j->t->f32};
  for(int64_t i = j->begin; i < j->end; ++i)
  {
    collect_ctx cc = {0, j->cands ? j->cands + j->offsets[i] : NULL};
    double q[6];
    load_query(j->prims, j->kind, j->t->dims, i, q);
    if(traverse(j->t, q, B, &fc, leaf_count, &cc, no_pref, NULL) < 0)
      j->overflow = 1;
    if(!j->cands) j->counts[i] = cc.count;
  }
  return NULL;
}

static void run_jobs(void* (*fn)(void*), void* jobs, size_t job_size, int nthreads)
{
  if(nthreads <= 1)
  {
    fn(jobs);
    return;
  }
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
  for(int k = 0; k < nthreads; ++k) pthread_create(&th[k], NULL, fn, (char*)jobs + k * job_size);
  for(int k = 0; k < nthreads; ++k) pthread_join(th[k], NULL);
  free(th);
}

/* LinearBVH.hpp:273-402: counts, exclusive-scan offsets, then candidates in
 * traversal (DFS) order per query. */
int64_t orc_bvh_find(const orc_bvh* t, int kind, const double* prims, int64_t nq, double tol,
                     int32_t* counts, int32_t* offsets, int32_t* cands, int nthreads)
{
  if(nthreads < 1) nthreads = 1;
  find_job* jobs = (find_job*)calloc((size_t)nthreads, sizeof(find_job));
  for(int k = 0; k < nthreads; ++k)
  {
    jobs[k].t = t;
    jobs[k].kind = kind;
    jobs[k].prims = prims;
    jobs[k].tol = tol;
    jobs[k].counts = counts;
    jobs[k].offsets = offsets;
    jobs[k].cands = cands;
    jobs[k].begin = nq * k / nthreads;
    jobs[k].end = nq * (k + 1) / nthreads;
  }
  int overflow = 0;
  int64_t total = 0;
  if(!cands)
  {
    run_jobs(find_worker, jobs, sizeof(find_job), nthreads);
    for(int k = 0; k < nthreads; ++k) overflow |= jobs[k].overflow;
    for(int64_t i = 0; i < nq; ++i)
    {
      offsets[i] = (int32_t)total;
      total += counts[i];
    }
  }
  else
  {
    run_jobs(find_worker, jobs, sizeof(find_job), nthreads);
    for(int k = 0; k < nthreads; ++k) overflow |= jobs[k].overflow;
    for(int64_t i = 0; i < nq; ++i) total += counts[i];
  }
  free(jobs);
  return overflow ? -1 : total;
}

int32_t orc_bvh_max_depth(const orc_bvh* t, int kind, const double* prims, int64_t nq,
                          double tol)
{
  pred_fn B = (kind == 0) ? pred_point : (kind == 1) ? pred_box : pred_ray;
  find_ctx fc = {tol, t->dims, t->f32};
  int32_t deepest = 0;
  for(int64_t i = 0; i < nq; ++i)
  {
    collect_ctx cc = {0, NULL};
    double q[6];
    load_query(prims, kind, t->dims, i, q);
    if(traverse(t, q, B, &fc, leaf_count, &cc, no_pref, &deepest) < 0) return -1;
  }
  return deepest;
}

/* Inner-node visits of the walk (bvh_traverse.hpp:66-154), summed over the
 * queries; -1 after a stack overflow. */
int64_t orc_bvh_visits(const orc_bvh* t, int kind, const double* prims, int64_t nq, double tol)
{
  pred_fn B = (kind == 0) ? pred_point : (kind == 1) ? pred_box : pred_ray;
  find_ctx fc = {tol, t->dims, t->f32};
  int64_t sum = 0;
  for(int64_t i = 0; i < nq; ++i)
  {
    collect_ctx cc = {0, NULL};
    double q[6];
    load_query(prims, kind, t->dims, i, q);
    const int32_t v = traverse(t, q, B, &fc, leaf_count, &cc, no_pref, NULL);
    if(v < 0) return -1;
    sum += v;
  }
  return sum;
}

int64_t orc_bvh_brute(const orc_bvh* t, int kind, const double* prims, int64_t nq, double tol,
                      int32_t* counts, int32_t* cands, int64_t cap)
{
  const int32_t n = t->n;
  /* scaled boxes in original-id order */
  orc_box* orig = (orc_box*)malloc(sizeof(orc_box) * n);
  for(int32_t p = 0; p < n; ++p) orig[t->leafs[p]] = t->leaf_aabbs[p];
  find_ctx fc = {tol, t->dims, t->f32};
  int64_t total = 0;
  for(int64_t i = 0; i < nq; ++i)
  {
    double pr[6];
    load_query(prims, kind, t->dims, i, pr);
    int32_t c = 0;
    for(int32_t b = 0; b < n; ++b)
    {
      if(!box_valid_d(&orig[b], t->dims)) continue;
      const int hit = (kind == 0) ? pred_point(pr, &orig[b], &fc)
        : (kind == 1)             ? pred_box(pr, &orig[b], &fc)
                                  : pred_ray(pr, &orig[b], &fc);
      if(hit)
      {
        if(cands && total < cap) cands[total] = b;
        total++;
        c++;
      }
    }
    counts[i] = c;
  }
  free(orig);
  return total;
}

/* ------------------------------------------------------------------------ */
/* SignedDistance: quest/SignedDistance.hpp:411-763                          */
/* ------------------------------------------------------------------------ */

struct orc_sdf
{
  int32_t nverts, ntris; /* ntris: cells */
  double *x, *y, *z;
  int32_t* conn;
  int32_t* off; /* cell offsets into conn (NULL: triangles, 3 nodes each) */
  int watertight, compute_sign;
  orc_box domain; /* m_boxDomain */
  orc_bvh* bvh;
};

/* off: ncells + 1 offsets into conn (cells of 3 or 4 nodes), or NULL for
 * triangles (SignedDistance.cpp:16-45 reads either mint mesh layout). */
static orc_sdf* sdf_new(const double* x, const double* y, const double* z, int32_t nverts,
                        const int32_t* conn, const int32_t* off, int32_t ntris, int watertight,
                        int compute_sign)
{
  orc_sdf* s = (orc_sdf*)calloc(1, sizeof(orc_sdf));
  s->nverts = nverts;
  s->ntris = ntris;
  const size_t nconn = off ? (size_t)off[ntris] : 3 * (size_t)ntris;
  s->x = (double*)malloc(sizeof(double) * nverts);
  s->y = (double*)malloc(sizeof(double) * nverts);
  s->z = (double*)malloc(sizeof(double) * nverts);
  s->conn = (int32_t*)malloc(sizeof(int32_t) * (nconn ? nconn : 1));
  memcpy(s->x, x, sizeof(double) * nverts);
  memcpy(s->y, y, sizeof(double) * nverts);
  memcpy(s->z, z, sizeof(double) * nverts);
  memcpy(s->conn, conn, sizeof(int32_t) * nconn);
  if(off)
  {
    s->off = (int32_t*)malloc(sizeof(int32_t) * ((size_t)ntris + 1));
    memcpy(s->off, off, sizeof(int32_t) * ((size_t)ntris + 1));
  }
  s->watertight = watertight;
  s->compute_sign = compute_sign;
  /* setMesh :470-475 (SEQ: addPoint over nodes) */
  box_clear(&s->domain);
  for(int32_t i = 0; i < nverts; ++i)
  {
    const double p[3] = {x[i], y[i], z[i]};
    box_add_point_d(&s->domain, p, 3);
  }
  /* getCellBoundingBox :609-633 */
  double* boxes = (double*)malloc(sizeof(orc_box) * (size_t)ntris);
  for(int32_t c = 0; c < ntris; ++c)
  {
    orc_box b;
    box_clear(&b);
    const int32_t b0 = s->off ? s->off[c] : 3 * c, b1 = s->off ? s->off[c + 1] : 3 * c + 3;
    for(int32_t k = b0; k < b1; ++k)
    {
      const int32_t v = s->conn[k];
      const double p[3] = {x[v], y[v], z[v]};
      box_add_point_d(&b, p, 3);
    }
    memcpy(boxes + 6 * (size_t)c, &b, sizeof(orc_box));
  }
  /* m_bvh.initialize(boxes, ncells) with the default scale 1.000123
   * (spin/BVH.hpp:410) */
  s->bvh = orc_bvh_new(boxes, ntris, 1.000123);
  free(boxes);
  return s;
}

orc_sdf* orc_sdf_new(const double* x, const double* y, const double* z, int32_t nverts,
                     const int32_t* conn, int32_t ntris, int watertight, int compute_sign)
{
  return sdf_new(x, y, z, nverts, conn, NULL, ntris, watertight, compute_sign);
}

orc_sdf* orc_sdf_new_cells(const double* x, const double* y, const double* z, int32_t nverts,
                           const int32_t* conn, const int32_t* off, int32_t ncells,
                           int watertight, int compute_sign)
{
  return sdf_new(x, y, z, nverts, conn, off, ncells, watertight, compute_sign);
}

void orc_sdf_free(orc_sdf* s)
{
  if(!s) return;
  free(s->x);
  free(s->y);
  free(s->z);
  free(s->conn);
  free(s->off);
  orc_bvh_free(s->bvh);
  free(s);
}

const orc_bvh* orc_sdf_bvh(const orc_sdf* s) { return s->bvh; }

void orc_sdf_bounds(const orc_sdf* s, double lo[3], double hi[3])
{
  memcpy(lo, s->domain.min, 3 * sizeof(double));
  memcpy(hi, s->domain.max, 3 * sizeof(double));
}

/* MinCandidate, SignedDistance.hpp:159-175. Location kinds follow
 * getClosestPointLocType :119-137: 0 vertex, 1 edge, 2 face, -1 unset. */
typedef struct
{
  double minSqDist;
  double minPt[3];
  int minType;
  int minLoc;
  int32_t minElem;
  int minSub; /* which triangle of the winning cell (minTri) */
  double sumN[3];
  int minCount;
  const orc_sdf* s;
  const double* q;
  int32_t leaves;
} min_cand;

static inline int loc_type(int loc) { return (loc < 0) ? 1 : (loc < 3) ? 0 : 2; }

/* Triangle sub of cell c: (n0, n1, n2), or (n0, n2, n3) for sub 1 of a quad
 * (SignedDistance.hpp:649-658). */
static inline int cell_tris(const orc_sdf* s, int32_t c)
{
  return (s->off && s->off[c + 1] - s->off[c] == 4) ? 2 : 1;
}
static inline void tri_points(const orc_sdf* s, int32_t c, int sub, double A[3], double B[3],
                              double C[3])
{
  const int32_t* n = s->conn + (s->off ? (size_t)s->off[c] : 3 * (size_t)c);
  const int32_t v[3] = {n[0], n[sub ? 2 : 1], n[sub ? 3 : 2]};
  A[0] = s->x[v[0]];
  A[1] = s->y[v[0]];
  A[2] = s->z[v[0]];
  B[0] = s->x[v[1]];
  B[1] = s->y[v[1]];
  B[2] = s->z[v[1]];
  C[0] = s->x[v[2]];
  C[1] = s->y[v[2]];
  C[2] = s->z[v[2]];
}

/* Triangle::normal, Triangle.hpp:99-102 */
static inline void tri_normal(const double A[3], const double B[3], const double C[3], double n[3])
{
  const double u[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
  const double v[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
  cross3(u, v, n);
}

/* Triangle::angle, Triangle.hpp:385-400 */
static inline double tri_angle(const double* P[3], int idx)
{
  const double* pt = P[idx];
  const double* p1 = P[(idx + 1) % 3];
  const double* p2 = P[(idx + 2) % 3];
  const double a[3] = {p1[0] - pt[0], p1[1] - pt[1], p1[2] - pt[2]};
  const double b[3] = {p2[0] - pt[0], p2[1] - pt[1], p2[2] - pt[2]};
  double ua[3], ub[3];
  orc_unit_vector(a, ua);
  orc_unit_vector(b, ub);
  double dp = dot3(ua, ub);
  dp = (dp < -1.0) ? -1.0 : (dp > 1.0) ? 1.0 : dp;
  return acos(dp);
}

/* checkCandidate, SignedDistance.hpp:637-737, for one triangle of the cell */
static void check_triangle(min_cand* m, int32_t cell, int sub)
{
  const double EPS = 1e-12;
  double A[3], B[3], C[3], cp[3];
  tri_points(m->s, cell, sub, A, B, C);
  const int loc = orc_closest_point(m->q, A, B, C, EPS, cp);
  const double sq = sqdist_pp(m->q, cp);
  const int type = loc_type(loc);
  const int shared = (type != 2);
  int upd = 0;
  const int cn = m->s->compute_sign;
  if(sq < m->minSqDist)
  {
    const int clear = !shared || (m->minType != type) || !near_eq(sqdist_pp(cp, m->minPt), 0., EPS);
    m->minSqDist = sq;
    memcpy(m->minPt, cp, sizeof(cp));
    m->minType = type;
    m->minLoc = loc;
    m->minElem = cell;
    m->minSub = sub;
    if(cn && clear)
    {
      m->sumN[0] = m->sumN[1] = m->sumN[2] = 0.0;
      m->minCount = 0;
    }
    upd = cn && shared;
  }
  else
  {
    upd = cn && shared && (m->minType == type) && near_eq(sqdist_pp(cp, m->minPt), 0., EPS);
  }
  if(upd)
  {
    ++m->minCount;
    double n[3], un[3];
    tri_normal(A, B, C, n);
    if(type == 1)
    {
      orc_unit_vector(n, un);
      for(int d = 0; d < 3; ++d) m->sumN[d] += un[d];
    }
    else if(type == 0)
    {
      /* Triangle::degenerate, Triangle.hpp:327-330: area = 0.5*|n| */
      const double area = 0.5 * sqrt(dot3(n, n));
      if(!near_eq(area, 0.0, 1.0e-12))
      {
        const double* P[3] = {A, B, C};
        const double alpha = tri_angle(P, loc);
        orc_unit_vector(n, un);
        for(int d = 0; d < 3; ++d) m->sumN[d] += un[d] * alpha;
      }
    }
  }
}

/* the cell's one or two triangles, in order */
static void check_candidate(min_cand* m, int32_t cell)
{
  const int nt = cell_tris(m->s, cell);
  for(int sub = 0; sub < nt; ++sub) check_triangle(m, cell, sub);
}

static void sdf_leaf(int32_t pos, const int32_t* leafs, void* ctx)
{
  min_cand* m = (min_cand*)ctx;
  m->leaves++;
  check_candidate(m, leafs[pos]);
}

/* traversePredicate, SignedDistance.hpp:576-578 */
static int sdf_pred(const void* prim, const orc_box* b, void* ctx)
{
  const min_cand* m = (const min_cand*)ctx;
  return orc_sqdist_point_box((const double*)prim, b) <= m->minSqDist;
}

/* getSurfaceNormal :741-749 */
static inline void surface_normal(const min_cand* m, double n[3])
{
  if(m->minType == 2)
  {
    double A[3], B[3], C[3];
    tri_points(m->s, m->minElem, m->minSub, A, B, C);
    tri_normal(A, B, C, n);
  }
  else
  {
    memcpy(n, m->sumN, 3 * sizeof(double));
  }
}

typedef struct
{
  const orc_sdf* s;
  const double* q;
  double *dist, *cp, *normal;
  int32_t *elem, *loc, *visits, *leaves;
  int64_t begin, end;
  int overflow;
} sdf_job;

static void* sdf_worker(void* arg)
{
  sdf_job* j = (sdf_job*)arg;
  const orc_sdf* s = j->s;
  for(int64_t i = j->begin; i < j->end; ++i)
  {
    const double* q = j->q + 3 * i;
    min_cand m;
    memset(&m, 0, sizeof(m));
    m.minSqDist = DBL_MAX;
    m.minType = -1;
    m.minLoc = 3;
    m.minElem = -1;
    m.s = s;
    m.q = q;
    const int32_t v = traverse(s->bvh, q, sdf_pred, &m, sdf_leaf, &m, point_pref, NULL);
    if(v < 0) j->overflow = 1;
    /* computeDistances :580-603 */
    double sgn = 1.0;
    if(s->compute_sign)
    {
      if(!(s->watertight && !box_contains(&s->domain, m.minPt)))
      {
        double n[3];
        surface_normal(&m, n);
        const double r[3] = {q[0] - m.minPt[0], q[1] - m.minPt[1], q[2] - m.minPt[2]};
        sgn = (dot3(r, n) >= 0.0) ? 1.0 : -1.0;
      }
    }
    j->dist[i] = sqrt(m.minSqDist) * sgn;
    if(j->cp) memcpy(j->cp + 3 * i, m.minPt, 3 * sizeof(double));
    if(j->normal)
    {
      double n[3];
      surface_normal(&m, n);
      orc_unit_vector(n, j->normal + 3 * i);
    }
    if(j->elem) j->elem[i] = m.minElem;
    if(j->loc) j->loc[i] = m.minLoc;
    if(j->visits) j->visits[i] = v;
    if(j->leaves) j->leaves[i] = m.leaves;
  }
  return NULL;
}

int32_t orc_sdf_max_depth(const orc_sdf* s, const double* q, int64_t nq)
{
  int32_t deepest = 0;
  for(int64_t i = 0; i < nq; ++i)
  {
    min_cand m;
    memset(&m, 0, sizeof(m));
    m.minSqDist = DBL_MAX;
    m.minType = -1;
    m.minLoc = 3;
    m.minElem = -1;
    m.s = s;
    m.q = q + 3 * i;
    if(traverse(s->bvh, q + 3 * i, sdf_pred, &m, sdf_leaf, &m, point_pref, &deepest) < 0)
      return -1;
  }
  return deepest;
}

int orc_sdf_eval(const orc_sdf* s, const double* q, int64_t nq, double* dist, double* cp,
                 double* normal, int32_t* elem, int32_t* loc, int32_t* visits, int32_t* leaves,
                 int nthreads)
{
  if(nthreads < 1) nthreads = 1;
  sdf_job* jobs = (sdf_job*)calloc((size_t)nthreads, sizeof(sdf_job));
  for(int k = 0; k < nthreads; ++k)
  {
    sdf_job* j = &jobs[k];
    j->s = s;
    j->q = q;
    j->dist = dist;
    j->cp = cp;
    j->normal = normal;
    j->elem = elem;
    j->loc = loc;
    j->visits = visits;
    j->leaves = leaves;
    j->begin = nq * k / nthreads;
    j->end = nq * (k + 1) / nthreads;
  }
  run_jobs(sdf_worker, jobs, sizeof(sdf_job), nthreads);
  int overflow = 0;
  for(int k = 0; k < nthreads; ++k) overflow |= jobs[k].overflow;
  free(jobs);
  return overflow ? -1 : 0;
}

typedef struct
{
  const orc_sdf* s;
  const double* q;
  double* sq;
  int32_t* elem;
  int64_t begin, end;
} brute_job;

static void* brute_worker(void* arg)
{
  brute_job* j = (brute_job*)arg;
  const orc_sdf* s = j->s;
  for(int64_t i = j->begin; i < j->end; ++i)
  {
    const double* q = j->q + 3 * i;
    double best = DBL_MAX;
    int32_t be = -1;
    for(int32_t c = 0; c < s->ntris; ++c)
      for(int sub = 0; sub < cell_tris(s, c); ++sub)
    {
      double A[3], B[3], C[3], cp[3];
      tri_points(s, c, sub, A, B, C);
      orc_closest_point(q, A, B, C, 1e-12, cp);
      const double d = sqdist_pp(q, cp);
      if(d < best)
      {
        best = d;
        be = c;
      }
    }
    j->sq[i] = best;
    if(j->elem) j->elem[i] = be;
  }
  return NULL;
}

int orc_sdf_brute(const orc_sdf* s, const double* q, int64_t nq, double* sqdist, int32_t* elem,
                  int nthreads)
{
  if(nthreads < 1) nthreads = 1;
  brute_job* jobs = (brute_job*)calloc((size_t)nthreads, sizeof(brute_job));
  for(int k = 0; k < nthreads; ++k)
  {
    jobs[k].s = s;
    jobs[k].q = q;
    jobs[k].sq = sqdist;
    jobs[k].elem = elem;
    jobs[k].begin = nq * k / nthreads;
    jobs[k].end = nq * (k + 1) / nthreads;
  }
  run_jobs(brute_worker, jobs, sizeof(brute_job), nthreads);
  free(jobs);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Adjoint (no reference counterpart).                                      */
/*   d(sd)/dq   = g * sgn * (q - cp)/|q - cp|                               */
/*   d(sd)/dV_k = -w_k * d(sd)/dq, w = barycentric weights of cp            */
/* Deterministic order: per triangle, per corner, queries in ascending index */
/* (partial sums from 0.0); per vertex, partials in ascending triangle id.  */
/* When q lies on the surface (|q-cp| = 0) the unit surface normal stands   */
/* in for the direction.                                                    */
/* ------------------------------------------------------------------------ */

/* Barycentric weights of the closest point for each location code of
 * closest_point.hpp:162-290 (recomputes the same v/w the forward uses). */
static void cp_weights(const double P[3], const double A[3], const double B[3], const double C[3],
                       double cp[3], double w[3], int* locp)
{
  const int loc = orc_closest_point(P, A, B, C, 1e-12, cp);
  double ab[3], ac[3], ap[3], bp[3], cpv[3];
  for(int d = 0; d < 3; ++d)
  {
    ab[d] = B[d] - A[d];
    ac[d] = C[d] - A[d];
    ap[d] = P[d] - A[d];
    bp[d] = P[d] - B[d];
    cpv[d] = P[d] - C[d];
  }
  const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  const double d5 = dot3(ab, cpv), d6 = dot3(ac, cpv);
  w[0] = w[1] = w[2] = 0.0;
  switch(loc)
  {
  case 0: w[0] = 1.0; break;
  case 1: w[1] = 1.0; break;
  case 2: w[2] = 1.0; break;
  case -1:
  {
    const double v = d1 / (d1 - d3);
    w[0] = 1.0 - v;
    w[1] = v;
    break;
  }
  case -3:
  {
    const double ww = d2 / (d2 - d6);
    w[0] = 1.0 - ww;
    w[2] = ww;
    break;
  }
  case -2:
  {
    const double ww = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    w[1] = 1.0 - ww;
    w[2] = ww;
    break;
  }
  default:
  {
    const double vc = d1 * d4 - d3 * d2;
    const double vb = d5 * d2 - d1 * d6;
    const double va = d3 * d6 - d5 * d4;
    const double denom = 1.0 / (va + vb + vc);
    const double v = vb * denom;
    const double ww = vc * denom;
    w[0] = 1.0 - v - ww;
    w[1] = v;
    w[2] = ww;
    break;
  }
  }
  *locp = loc;
}

static int cmp_pair(const void* a, const void* b)
{
  const int64_t* x = (const int64_t*)a;
  const int64_t* y = (const int64_t*)b;
  return (x[0] < y[0]) ? -1 : (x[0] > y[0]) ? 1 : (x[1] < y[1]) ? -1 : (x[1] > y[1]);
}

int orc_sdf_backward(const orc_sdf* s, const double* q, int64_t nq, const double* dist,
                     const int32_t* elem, const float* gout, double* grad_q, double* grad_v)
{
  if(s->off) return -1; /* the adjoint is for triangle meshes */
  const int32_t nt = s->ntris, nv = s->nverts;
  /* elem -1: no closest triangle (gradient 0); any other id out of range fails */
  for(int64_t i = 0; i < nq; ++i)
    if(elem[i] < -1 || elem[i] >= nt) return -1;
  /* per-query direction u = sgn*(q-cp)/|q-cp| scaled by g */
  double* dir = (double*)malloc(sizeof(double) * 3 * (size_t)nq);
  double* wts = (double*)malloc(sizeof(double) * 3 * (size_t)nq);
  for(int64_t i = 0; i < nq; ++i)
  {
    const double* P = q + 3 * i;
    const int32_t c = elem[i];
    if(c < 0)
    {
      for(int d = 0; d < 3; ++d)
      {
        dir[3 * i + d] = 0.0;
        wts[3 * i + d] = 0.0;
        if(grad_q) grad_q[3 * i + d] = 0.0;
      }
      continue;
    }
    double A[3], B[3], C[3], cp[3], w[3];
    int loc;
    tri_points(s, c, 0, A, B, C);  /* the adjoint is for triangle meshes */
    cp_weights(P, A, B, C, cp, w, &loc);
    const double sgn = signbit(dist[i]) ? -1.0 : 1.0;
    double r[3] = {P[0] - cp[0], P[1] - cp[1], P[2] - cp[2]};
    const double len = sqrt(dot3(r, r));
    double u[3];
    if(len > 0.0)
    {
      const double inv = 1.0 / len;
      for(int d = 0; d < 3; ++d) u[d] = r[d] * inv;
    }
    else
    {
      double n[3];
      tri_normal(A, B, C, n);
      orc_unit_vector(n, u);
    }
    const double g = (double)gout[i] * sgn;
    for(int d = 0; d < 3; ++d)
    {
      dir[3 * i + d] = g * u[d];
      wts[3 * i + d] = w[d];
    }
    if(grad_q)
      for(int d = 0; d < 3; ++d) grad_q[3 * i + d] = dir[3 * i + d];
  }
  if(grad_v)
  {
    /* queries grouped by triangle, ascending query index */
    int64_t* pairs = (int64_t*)malloc(sizeof(int64_t) * 2 * (size_t)nq);
    for(int64_t i = 0; i < nq; ++i)
    {
      pairs[2 * i] = elem[i];
      pairs[2 * i + 1] = i;
    }
    qsort(pairs, (size_t)nq, 2 * sizeof(int64_t), cmp_pair);
    double* part = (double*)calloc((size_t)nt * 9, sizeof(double));
    for(int64_t k = 0; k < nq; ++k)
    {
      const int64_t c = pairs[2 * k], i = pairs[2 * k + 1];
      if(c < 0) continue;
      for(int corner = 0; corner < 3; ++corner)
        for(int d = 0; d < 3; ++d)
          part[9 * c + 3 * corner + d] += -(wts[3 * i + corner] * dir[3 * i + d]);
    }
    for(int64_t v = 0; v < 3 * (int64_t)nv; ++v) grad_v[v] = 0.0;
    for(int32_t c = 0; c < nt; ++c)
      for(int corner = 0; corner < 3; ++corner)
      {
        const int32_t v = s->conn[3 * (size_t)c + corner];
        for(int d = 0; d < 3; ++d) grad_v[3 * (size_t)v + d] += part[9 * (size_t)c + 3 * corner + d];
      }
    free(part);
    free(pairs);
  }
  free(dir);
  free(wts);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Triangle-triangle intersection (Devillers & Guigue 2002, as the reference */
/* states it) and the BVH mesh self-intersection test.                       */
/* ------------------------------------------------------------------------ */

/* twoDcross, primal/operators/detail/intersect_impl.hpp:538-544 */
static inline double cross2(const double a[2], const double b[2], const double c[2])
{
  return ((a[0] - c[0]) * (b[1] - c[1]) - (a[1] - c[1]) * (b[0] - c[0]));
}
/* isGpeq / isLpeq against 0, fuzzy_comparators.hpp:55-84 */
static inline int gpeq0(double x, int inc, double eps)
{
  return (inc && near_eq(x, 0.0, eps)) || is_gt(x, 0.0, eps);
}
static inline int lpeq0(double x, int inc, double eps)
{
  return (inc && near_eq(x, 0.0, eps)) || is_lt(x, 0.0, eps);
}

/* checkEdge, intersect_impl.hpp:1344-1392 */
static int tt_check_edge(const double* p1, const double* q1, const double* r1, const double* p2,
                         const double* r2, int inc, double eps)
{
  if(gpeq0(cross2(r2, p2, q1), inc, eps))
  {
    if(!gpeq0(cross2(r2, p1, q1), inc, eps)) return 0;
    if(gpeq0(cross2(p1, p2, q1), inc, eps)) return 1;
    return gpeq0(cross2(p1, p2, r1), inc, eps) && gpeq0(cross2(q1, r1, p2), inc, eps);
  }
  return gpeq0(cross2(r2, p2, r1), inc, eps) && gpeq0(cross2(q1, r1, r2), inc, eps) &&
    gpeq0(cross2(p1, p2, r1), inc, eps);
}

/* checkVertex, intersect_impl.hpp:1394-1484 */
static int tt_check_vertex(const double* p1, const double* q1, const double* r1, const double* p2,
                           const double* q2, const double* r2, int inc, double eps)
{
  if(gpeq0(cross2(r2, p2, q1), inc, eps))
  {
    if(gpeq0(cross2(q2, r2, q1), inc, eps))
    {
      if(gpeq0(cross2(p1, p2, q1), inc, eps)) return lpeq0(cross2(p1, q2, q1), inc, eps);
      return gpeq0(cross2(p1, p2, r1), inc, eps) && lpeq0(cross2(q1, p2, r1), inc, eps);
    }
    return lpeq0(cross2(p1, q2, q1), inc, eps) && gpeq0(cross2(q2, r2, r1), inc, eps) &&
      gpeq0(cross2(q1, r1, q2), inc, eps);
  }
  if(!gpeq0(cross2(r2, p2, r1), inc, eps)) return 0;
  if(gpeq0(cross2(q1, r1, r2), inc, eps)) return gpeq0(cross2(r1, p1, p2), inc, eps);
  return gpeq0(cross2(q1, r1, q2), inc, eps) && gpeq0(cross2(q2, r2, r1), inc, eps);
}

/* intersectPermuted2DTriangles, intersect_impl.hpp:1281-1342: where p1 lies
 * against the three edges of triangle 2 picks the edge or vertex test. */
static int tt_permuted_2d(const double* p1, const double* q1, const double* r1, const double* p2,
                          const double* q2, const double* r2, int inc, double eps)
{
  const int a = gpeq0(cross2(p2, q2, p1), inc, eps);
  const int b = gpeq0(cross2(q2, r2, p1), inc, eps);
  const int c = gpeq0(cross2(r2, p2, p1), inc, eps);
  if(a && b) return c ? 1 : tt_check_edge(p1, q1, r1, p2, r2, inc, eps);
  if(a) return c ? tt_check_edge(p1, q1, r1, r2, q2, inc, eps)
                 : tt_check_vertex(p1, q1, r1, p2, q2, r2, inc, eps);
  if(b) return c ? tt_check_edge(p1, q1, r1, q2, p2, inc, eps)
                 : tt_check_vertex(p1, q1, r1, q2, r2, p2, inc, eps);
  return tt_check_vertex(p1, q1, r1, r2, p2, q2, inc, eps);
}

/* TriangleIntersection2D, intersect_impl.hpp:1251-1279: both triangles made
 * counter-clockwise by swapping their last two corners. */
static int tt_tri_2d(const double t1[6], const double t2[6], int inc, double eps)
{
  const int f1 = is_lt(cross2(t1, t1 + 2, t1 + 4), 0.0, eps);
  const int f2 = is_lt(cross2(t2, t2 + 2, t2 + 4), 0.0, eps);
  return tt_permuted_2d(t1, f1 ? t1 + 4 : t1 + 2, f1 ? t1 + 2 : t1 + 4, t2, f2 ? t2 + 4 : t2 + 2,
                        f2 ? t2 + 2 : t2 + 4, inc, eps);
}

/* intersectCoplanar3DTriangles, intersect_impl.hpp:1190-1249: project on the
 * coordinate plane where triangle 1's normal is largest. */
static int tt_coplanar(const double* p1, const double* q1, const double* r1, const double* p2,
                       const double* q2, const double* r2, const double n_in[3], int inc,
                       double eps)
{
  const double n[3] = {fabs(n_in[0]), fabs(n_in[1]), fabs(n_in[2])};
  int u, v, qpr; /* coordinates kept; qpr: corner order q,p,r instead of p,q,r */
  if(is_gt(n[0], n[2], eps) && is_geq(n[0], n[1], eps))
  {
    u = 2;
    v = 1;
    qpr = 1;
  }
  else if(is_gt(n[1], n[2], eps) && is_geq(n[1], n[0], eps))
  {
    u = 0;
    v = 2;
    qpr = 1;
  }
  else
  {
    u = 0;
    v = 1;
    qpr = 0;
  }
  const double* c1[3] = {qpr ? q1 : p1, qpr ? p1 : q1, r1};
  const double* c2[3] = {qpr ? q2 : p2, qpr ? p2 : q2, r2};
  double a[6], b[6];
  for(int k = 0; k < 3; ++k)
  {
    a[2 * k] = c1[k][u];
    a[2 * k + 1] = c1[k][v];
    b[2 * k] = c2[k][u];
    b[2 * k + 1] = c2[k][v];
  }
  return tt_tri_2d(a, b, inc, eps);
}

/* intersectTwoPermutedTriangles, intersect_impl.hpp:456-471 (Devillers &
 * Guigue's interval test): Vector3(a,b) = b - a, Triangle3(a,b,c).normal() =
 * (b - a) x (c - a). */
static int tt_two_permuted(const double* p1, const double* q1, const double* r1,
                           const double* p2, const double* q2, const double* r2, int inc,
                           double eps)
{
  double u[3], v[3], w[3], n[3];
  for(int d = 0; d < 3; ++d)
  {
    u[d] = q2[d] - q1[d];
    v[d] = p2[d] - q1[d];
    w[d] = p1[d] - q1[d];
  }
  cross3(v, w, n);
  if(!lpeq0(dot3(u, n), inc, eps)) return 0;
  for(int d = 0; d < 3; ++d)
  {
    u[d] = r2[d] - p1[d];
    v[d] = p2[d] - p1[d];
    w[d] = r1[d] - p1[d];
  }
  cross3(v, w, n);
  return lpeq0(dot3(u, n), inc, eps);
}

/* The permutation step of intersect_tri3D_tri3D (intersect_impl.hpp:213-381
 * for triangle 1) and of intersectOnePermutedTriangle (:1116-1187 for
 * triangle 2), which share one table: from the signed distances d of one
 * triangle's corners to the other's plane, the rotation that brings the lone
 * corner first (rot: corner rot becomes p) and whether the OTHER triangle's
 * last two corners swap. Returns -1 when all three are near zero. */
static int tt_orient(const double d[3], double eps, int* rot, int* swap)
{
  const double p = d[0], q = d[1], r = d[2];
  if(is_gt(p, 0.0, eps))
  {
    if(is_gt(q, 0.0, eps)) { *rot = 2; *swap = 1; }
    else if(is_gt(r, 0.0, eps)) { *rot = 1; *swap = 1; }
    else { *rot = 0; *swap = 0; }
  }
  else if(is_lt(p, 0.0, eps))
  {
    if(is_lt(q, 0.0, eps)) { *rot = 2; *swap = 0; }
    else if(is_lt(r, 0.0, eps)) { *rot = 1; *swap = 0; }
    else { *rot = 0; *swap = 1; }
  }
  else if(is_lt(q, 0.0, eps))
  {
    if(is_geq(r, 0.0, eps)) { *rot = 1; *swap = 1; }
    else { *rot = 0; *swap = 0; }
  }
  else if(is_gt(q, 0.0, eps))
  {
    if(is_gt(r, 0.0, eps)) { *rot = 0; *swap = 1; }
    else { *rot = 1; *swap = 0; }
  }
  else if(is_gt(r, 0.0, eps)) { *rot = 2; *swap = 0; }
  else if(is_lt(r, 0.0, eps)) { *rot = 2; *swap = 1; }
  else return -1;
  return 0;
}

/* nonzeroSignMatch / countZeros / twoZeros / oneZeroOthersMatch,
 * intersect_impl.hpp:547-592 */
static inline int sgn(double x) { return (0 < x) - (x < 0); }
static int tt_all_one_side(const double d[3], double eps)
{
  return !near_eq(d[0], 0., eps) && !near_eq(d[1], 0., eps) && !near_eq(d[2], 0., eps) &&
    sgn(d[0]) == sgn(d[1]) && sgn(d[0]) == sgn(d[2]);
}
static int tt_touch_only(const double d[3], double eps)
{
  const int z0 = near_eq(d[0], 0.0, eps), z1 = near_eq(d[1], 0.0, eps), z2 = near_eq(d[2], 0.0, eps);
  const int nz = z0 + z1 + z2;
  if(nz == 2) return 1;
  return nz == 1 &&
    ((z0 && is_gt(d[1] * d[2], 0.0, eps)) || (z1 && is_gt(d[2] * d[0], 0.0, eps)) ||
     (z2 && is_gt(d[0] * d[1], 0.0, eps)));
}

/* Triangle::normal().unitVector() (Triangle.hpp:99-102, Vector.hpp:477-490)
 * and the signed distances (t[k] - o[2]) . n of triangle t's corners. */
static void tt_plane_dists(const double* t, const double* o, double n[3], double d[3])
{
  double u[3], v[3], c[3];
  for(int k = 0; k < 3; ++k)
  {
    u[k] = o[3 + k] - o[k];
    v[k] = o[6 + k] - o[k];
  }
  cross3(u, v, c);
  orc_unit_vector(c, n);
  for(int k = 0; k < 3; ++k)
  {
    const double e[3] = {t[3 * k] - o[6], t[3 * k + 1] - o[7], t[3 * k + 2] - o[8]};
    d[k] = dot3(e, n);
  }
}

/* primal::intersect(Triangle3, Triangle3, includeBoundary, EPS)
 * (primal/operators/intersect.hpp:65-71 -> intersect_impl.hpp:155-382). */
int orc_tri_tri(const double t1[9], const double t2[9], int include_boundary, double eps)
{
  const int inc = include_boundary != 0;
  double n1[3], n2[3], d1[3], d2[3];
  tt_plane_dists(t1, t2, n2, d1);
  if(tt_all_one_side(d1, eps)) return 0;
  if(!inc && tt_touch_only(d1, eps)) return 0;
  tt_plane_dists(t2, t1, n1, d2);
  if(tt_all_one_side(d2, eps)) return 0;
  if(!inc && tt_touch_only(d2, eps)) return 0;
  int rot1, swap2;
  if(tt_orient(d1, eps, &rot1, &swap2) < 0)
    return tt_coplanar(t1, t1 + 3, t1 + 6, t2, t2 + 3, t2 + 6, n1, inc, eps);
  const double* P1 = t1 + 3 * rot1;
  const double* Q1 = t1 + 3 * ((rot1 + 1) % 3);
  const double* R1 = t1 + 3 * ((rot1 + 2) % 3);
  const double* c2[3] = {t2, swap2 ? t2 + 6 : t2 + 3, swap2 ? t2 + 3 : t2 + 6};
  const double e2[3] = {d2[0], swap2 ? d2[2] : d2[1], swap2 ? d2[1] : d2[2]};
  int rot2, swap1;
  if(tt_orient(e2, eps, &rot2, &swap1) < 0)
    return tt_coplanar(P1, Q1, R1, c2[0], c2[1], c2[2], n1, inc, eps);
  return tt_two_permuted(P1, swap1 ? R1 : Q1, swap1 ? Q1 : R1, c2[rot2], c2[(rot2 + 1) % 3],
                         c2[(rot2 + 2) % 3], inc, eps);
}

/* quest::findTriMeshIntersectionsBVH (quest/MeshTester.hpp:66-87) through
 * CandidateFinderBase::initialize / findTriMeshIntersections and the BVH
 * CandidateFinder (quest/detail/MeshTester_detail.hpp:159-341): per triangle
 * its corner box and the degenerate flag (Triangle::degenerate, Triangle.hpp:
 * 104-109,327-330: |0.5 |n|| <= 1e-12); a BVH (default scale 1.000123) over
 * the boxes; findBoundingBoxes of every triangle's box; the pairs (i, j) with
 * i < j in candidate order, kept when the triangles intersect (boundary
 * excluded, tolerance thresh). Writes up to cap pairs in the sequential
 * order and degen[ntris] flags; returns the number of pairs, or -1. */
int64_t orc_mesh_self_isect(const double* x, const double* y, const double* z, int32_t nverts,
                            const int32_t* conn, int32_t ntris, double thresh, int32_t* first,
                            int32_t* second, int64_t cap, uint8_t* degen, int nthreads)
{
  (void)nverts;
  double* tri = (double*)malloc(sizeof(double) * 9 * (size_t)(ntris > 0 ? ntris : 1));
  double* boxes = (double*)calloc(6 * (size_t)(ntris > 0 ? ntris : 1), sizeof(double));
  for(int32_t c = 0; c < ntris; ++c)
  {
    double* t = tri + 9 * (size_t)c;
    for(int k = 0; k < 3; ++k)
    {
      const int32_t v = conn[3 * (size_t)c + k];
      t[3 * k] = x[v];
      t[3 * k + 1] = y[v];
      t[3 * k + 2] = z[v];
    }
    double u[3], w[3], n[3];
    for(int k = 0; k < 3; ++k)
    {
      u[k] = t[3 + k] - t[k];
      w[k] = t[6 + k] - t[k];
    }
    cross3(u, w, n);
    degen[c] = (uint8_t)near_eq(0.5 * sqrt(dot3(n, n)), 0.0, 1.0e-12);
    /* BoundingBox{A, B, C} (compute_bounding_box.hpp:110-114) */
    double* b = boxes + 6 * (size_t)c;
    for(int k = 0; k < 3; ++k) b[k] = b[3 + k] = t[k];
    for(int corner = 1; corner < 3; ++corner)
      for(int k = 0; k < 3; ++k)
      {
        const double p = t[3 * corner + k];
        if(p < b[k]) b[k] = p;
        if(p > b[3 + k]) b[3 + k] = p;
      }
  }
  orc_bvh* bvh = orc_bvh_new(boxes, ntris, 1.000123);
  int32_t* counts = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ntris > 0 ? ntris : 1));
  int32_t* offsets = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ntris > 0 ? ntris : 1));
  int64_t total = orc_bvh_find(bvh, 1, boxes, ntris, 0.0, counts, offsets, NULL, nthreads);
  int64_t npairs = -1;
  if(total >= 0)
  {
    int32_t* cands = (int32_t*)malloc(sizeof(int32_t) * (size_t)(total > 0 ? total : 1));
    orc_bvh_find(bvh, 1, boxes, ntris, 0.0, counts, offsets, cands, nthreads);
    npairs = 0;
    for(int32_t i = 0; i < ntris; ++i)
      for(int32_t k = 0; k < counts[i]; ++k)
      {
        const int32_t j = cands[offsets[i] + k];
        if(i < j && orc_tri_tri(tri + 9 * (size_t)i, tri + 9 * (size_t)j, 0, thresh))
        {
          if(npairs < cap)
          {
            first[npairs] = i;
            second[npairs] = j;
          }
          ++npairs;
        }
      }
    free(cands);
  }
  free(offsets);
  free(counts);
  orc_bvh_free(bvh);
  free(boxes);
  free(tri);
  return npairs;
}
