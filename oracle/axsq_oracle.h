/*
 * axsq_oracle.h -- CPU restatement of Axom's spatial-query path (TEST INFRASTRUCTURE).
 *
 * This library is the parity checker for the HIP implementation in axom_amd/.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the checker or as the timed CPU baseline. The product path
 * never calls it.
 *
 * It restates, in plain C, the sequential (SEQ_EXEC, no RAJA) path of:
 *   - spin::BVH build         src/axom/spin/internal/linear_bvh/build_radix_tree.hpp
 *                             src/axom/spin/policy/LinearBVH.hpp:193-269
 *   - spin::BVH traversal     src/axom/spin/internal/linear_bvh/bvh_traverse.hpp:66-154
 *                             src/axom/spin/policy/LinearBVH.hpp:57-109,273-402
 *   - quest::SignedDistance   src/axom/quest/SignedDistance.hpp:411-763
 * plus primal predicates (BoundingBox, closest_point, squared_distance,
 * intersect_ray) and an adjoint that has no reference counterpart.
 *
 * Build flags must keep -ffp-contract=off and no fast-math so that every
 * double operation rounds exactly like the reference's.
 */
#ifndef AXSQ_ORACLE_H
#define AXSQ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Layout of primal::BoundingBox<double,3>: min xyz then max xyz (48 B). */
typedef struct orc_box
{
  double min[3];
  double max[3];
} orc_box;

typedef struct orc_bvh
{
  int32_t n;            /* leaves after padding a 0/1-box input to 2 */
  orc_box bounds;       /* union of the scaled leaf boxes (BVH::getBounds) */
  uint32_t* mcodes;     /* n Morton codes in sorted order */
  int32_t* leafs;       /* n: sorted position -> original box id */
  int32_t* lchild;      /* n-1 radix-tree children (>= n-1 means leaf) */
  int32_t* rchild;      /* n-1 */
  int32_t* parents;     /* 2n-1 */
  orc_box* inner_aabbs; /* n-1 */
  orc_box* leaf_aabbs;  /* n scaled leaf boxes in sorted order */
  orc_box* nodes;       /* 2(n-1): child boxes of inner node i at 2i, 2i+1 */
  int32_t* children;    /* 2(n-1): 2*child for inner, -(pos+1) for leaf */
  int32_t dims;         /* 3, or 2 for BVH<2>: boxes use only their x, y entries */
  int32_t f32;          /* a float tree (FloatType = float): float values in the boxes */
} orc_bvh;

/* ---- primal restatements, exported for unit tests ------------------------ */
uint32_t orc_morton3(int32_t x, int32_t y, int32_t z);
uint32_t orc_morton2(int32_t x, int32_t y);
void orc_box_scale(orc_box* b, double sf);
int orc_box_valid(const orc_box* b);
void orc_box_add_box(orc_box* b, const orc_box* o);
double orc_sqdist_point_box(const double p[3], const orc_box* b);
int orc_ray_box(const double o[3], const double d[3], const orc_box* b, double eps);
int orc_closest_point(const double p[3], const double a[3], const double b[3],
                      const double c[3], double eps, double cp[3]);
void orc_unit_vector(const double v[3], double out[3]);
void orc_ray_normalize(const double d[3], double out[3]);

/* ---- BVH ---------------------------------------------------------------- */
orc_bvh* orc_bvh_new(const double* boxes /* AoS n x 6 */, int32_t n, double scale);
/* spin::BVH<2>: boxes AoS n x 4 (min x y, max x y); queries then are 2-D
 * (points 2 doubles, boxes and rays 4) */
orc_bvh* orc_bvh_new2d(const double* boxes /* AoS n x 4 */, int32_t n, double scale);
/* A float tree (spin::BVH<nd, Exec, float>): boxes AoS n x 2nd floats, the
 * scale, bounds, Morton normalisation and ray test in float. Queries are
 * passed as doubles holding float values. */
orc_bvh* orc_bvh_new_f32(const float* boxes, int32_t n, int nd, float scale);
int orc_ray_box_f32(const double o[3], const double dir[3], const orc_box* b, float eps);
void orc_bvh_free(orc_bvh* t);
int32_t orc_bvh_size(const orc_bvh* t);
void orc_bvh_export(const orc_bvh* t, double* bounds6, uint32_t* mcodes,
                    int32_t* leafs, int32_t* lchild, int32_t* rchild,
                    int32_t* parents, double* inner_aabbs, double* nodes,
                    int32_t* children);

/* kind: 0 = points (box contains point), 1 = boxes (intersectsWith),
 *       2 = rays (slab test with tolerance tol; direction already unit).
 * prims is AoS: 3 doubles per point, 6 per box (min,max) or ray (origin,dir);
 * for a 2-D tree 2 per point and 4 per box or ray.
 * With cands == NULL only counts/offsets are written; returns the total.   */
int64_t orc_bvh_find(const orc_bvh* t, int kind, const double* prims, int64_t nq,
                     double tol, int32_t* counts, int32_t* offsets, int32_t* cands,
                     int nthreads);
/* The deepest traversal stack entry (1 = one pending node) any of the nq
 * queries' walks pushes; -1 when a walk overflows the 64-entry stack. */
int32_t orc_bvh_max_depth(const orc_bvh* t, int kind, const double* prims, int64_t nq,
                          double tol);
/* Inner-node visits of the nq queries' walks, summed; -1 on an overflow. */
int64_t orc_bvh_visits(const orc_bvh* t, int kind, const double* prims, int64_t nq, double tol);
/* O(N*Q) brute force over the scaled leaf boxes, ascending box id. */
int64_t orc_bvh_brute(const orc_bvh* t, int kind, const double* prims, int64_t nq,
                      double tol, int32_t* counts, int32_t* cands, int64_t cap);

/* ---- SignedDistance ----------------------------------------------------- */
typedef struct orc_sdf orc_sdf;
orc_sdf* orc_sdf_new(const double* x, const double* y, const double* z, int32_t nverts,
                     const int32_t* conn, int32_t ntris, int watertight, int compute_sign);
/* A mesh of triangle and quad cells: off = ncells + 1 offsets into conn
 * (3 or 4 nodes per cell); a quad is checked as (n0,n1,n2) then (n0,n2,n3). */
orc_sdf* orc_sdf_new_cells(const double* x, const double* y, const double* z, int32_t nverts,
                           const int32_t* conn, const int32_t* off, int32_t ncells,
                           int watertight, int compute_sign);
void orc_sdf_free(orc_sdf* s);
const orc_bvh* orc_sdf_bvh(const orc_sdf* s);
void orc_sdf_bounds(const orc_sdf* s, double lo[3], double hi[3]);
/* q is AoS 3*nq. Optional outputs may be NULL. elem = winning triangle id,
 * loc = closest_point location code, visits/leaves = per-query inner-node
 * visits and leaf checks of the reference traversal.                       */
int orc_sdf_eval(const orc_sdf* s, const double* q, int64_t nq, double* dist,
                 double* cp, double* normal, int32_t* elem, int32_t* loc,
                 int32_t* visits, int32_t* leaves, int nthreads);
/* The deepest stack entry any of the nq signed-distance walks pushes (-1 on
 * overflow). */
int32_t orc_sdf_max_depth(const orc_sdf* s, const double* q, int64_t nq);
/* Brute-force min over all triangles of |q - closest_point(q,T)|^2. */
int orc_sdf_brute(const orc_sdf* s, const double* q, int64_t nq, double* sqdist,
                  int32_t* elem, int nthreads);
/* Adjoint of the signed distance (no reference counterpart). Recomputes the
 * closest point on elem[i]; grad_q is AoS 3*nq, grad_v AoS 3*nverts.        */
int orc_sdf_backward(const orc_sdf* s, const double* q, int64_t nq, const double* dist,
                     const int32_t* elem, const float* gout, double* grad_q,
                     double* grad_v);

/* primal::intersect(Triangle3, Triangle3, includeBoundary, EPS):
 * t1, t2 = corners A, B, C (9 doubles each); 1 if they intersect. */
int orc_tri_tri(const double t1[9], const double t2[9], int include_boundary, double eps);
/* quest::findTriMeshIntersectionsBVH: pairs in sequential order (up to cap),
 * degenerate flags; returns the number of pairs or -1. */
int64_t orc_mesh_self_isect(const double* x, const double* y, const double* z, int32_t nverts,
                            const int32_t* conn, int32_t ntris, double thresh, int32_t* first,
                            int32_t* second, int64_t cap, uint8_t* degen, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
