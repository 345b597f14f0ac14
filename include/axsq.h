/*
 * axsq.h -- C-ABI of the MI355X (gfx950) spatial-query library libaxsq.so.
 *
 * This is the drop-in boundary for Axom's spatial-query hot path:
 *   spin::BVH<3, ExecSpace>::initialize / findPoints / findBoundingBoxes /
 *   findRays / getBounds, and quest::SignedDistance<3, ExecSpace>
 *   construction + computeDistances, plus the signed-distance adjoint.
 *
 * Conventions
 *   - Every pointer to bulk data is a DEVICE pointer (hipMalloc'd or managed),
 *     the same contract the reference states for GPU exec spaces
 *     (src/axom/spin/BVH.hpp:198-203,238-241).
 *   - Boxes are primal::BoundingBox<double,3> laid out as 6 doubles
 *     {min x,y,z, max x,y,z} (48 B, the reference's AoS layout).
 *   - Point queries are given as three coordinate pointers and an element
 *     stride in doubles: AoS Point<double,3> arrays pass (p, p+1, p+2, 3),
 *     SoA ZipPoint-style arrays pass (x, y, z, 1).
 *   - Every call returns AXSQ_OK (0) or a negative status; the text of the
 *     last error on the calling thread is returned by axsq_last_error().
 *     Nothing aborts (the reference aborts through SLIC_ERROR instead).
 *   - Calls are stream-ordered on `stream` (a hipStream_t, NULL = default
 *     stream) and return after their results are complete, matching the
 *     reference's synchronous HIP_EXEC (core/execution/internal/hip_exec.hpp:38-85).
 *   - IndexType is int32 (core/Types.hpp:63-65 default).
 */
#ifndef AXSQ_H
#define AXSQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AXSQ_OK 0
#define AXSQ_ERR_INVALID_ARG (-1)
#define AXSQ_ERR_HIP (-2)
#define AXSQ_ERR_NOMEM (-3)
#define AXSQ_ERR_STACK_OVERFLOW (-4)
#define AXSQ_ERR_STATE (-5)

/* Precision of the tree traversal. AXSQ_FP64 is the reference's plain form;
 * AXSQ_FP32_TRAVERSAL walks FP32 nodes whose boxes are rounded outward and
 * finishes every leaf test in FP64. */
#define AXSQ_FP64 0
#define AXSQ_FP32_TRAVERSAL 1

/* Query kinds of axsq_bvh_find*. */
#define AXSQ_QUERY_POINTS 0 /* BoundingBox::contains(Point), BVH.hpp:499-501 */
#define AXSQ_QUERY_BOXES 1  /* BoundingBox::intersectsWith, BVH.hpp:558-560 */
#define AXSQ_QUERY_RAYS 2   /* intersect_ray(Ray, Box, ip, tol), BVH.hpp:529-532 */

typedef struct axsq_bvh axsq_bvh;
typedef struct axsq_sdf axsq_sdf;

/* Memory that the candidate arrays of axsq_bvh_find are allocated in: the
 * allocator id a caller gives BVH::setAllocatorID (src/axom/spin/BVH.hpp:
 * 251-261; the candidate array is allocated with it at
 * src/axom/spin/policy/LinearBVH.hpp:338-341). */
#define AXSQ_ALLOC_DEVICE 0      /* hipMalloc: device memory (the default) */
#define AXSQ_ALLOC_MANAGED 1     /* hipMallocManaged: readable on host and device */
#define AXSQ_ALLOC_PINNED_HOST 2 /* hipHostMalloc: host memory the device writes */

/* Text of the last error raised on the calling thread ("" if none). */
const char* axsq_last_error(void);
/* Library version as 10000*major + 100*minor + patch. */
int axsq_version(void);
/* Frees memory returned by the library (candidate arrays, pair lists), in
 * whichever of the AXSQ_ALLOC_* memories it was allocated. */
int axsq_free(void* ptr);
/* Frees host memory returned by the library (axsq_stl_read). */
int axsq_host_free(void* host_ptr);
/* Frees the calling thread's scratch memory on the current device: the
 * sort, scan and candidate-search buffers, the signed-distance walk records
 * and tables, the mesh tester's tree. Later calls allocate again as needed.
 * Must not run while work of this thread that uses them is in flight. */
int axsq_release_workspace(void);

/* ---- BVH ----------------------------------------------------------------
 * Builds a linear BVH over n boxes.
 * Replaces spin::BVH<3,Exec,double>::initialize (src/axom/spin/BVH.hpp:426-477)
 * -> LinearBVH::buildImpl (src/axom/spin/policy/LinearBVH.hpp:193-269)
 * -> build_radix_tree (src/axom/spin/internal/linear_bvh/build_radix_tree.hpp:579-611).
 * scale: BVH::setScaleFactor (BVH.hpp:269, default 1.000123).
 * fp_mode: AXSQ_FP64, or AXSQ_FP32_TRAVERSAL to also emit FP32 nodes.
 * n may be 0 or 1 (padded to 2 leaves with an invalid box, BVH.hpp:441-464). */
int axsq_bvh_build(axsq_bvh** out, const double* boxes, int64_t n, double scale,
                   int fp_mode, void* stream);
int axsq_bvh_destroy(axsq_bvh* bvh);

/* Builds a 2-D linear BVH over n boxes, AoS of 4 doubles (min x y, max x y).
 * Replaces spin::BVH<2,Exec,double>::initialize (BVH.hpp:426-477 with
 * NDIMS = 2): the same pipeline with 16-bit-per-axis Morton codes
 * (build_radix_tree.hpp:48-65 with Dims = 2). The handle then serves
 * axsq_bvh_find* with 2-D queries -- points x, y (z ignored, may be NULL),
 * boxes AoS of 4 doubles, rays {origin x y, unit direction x y}, strides in
 * doubles -- and axsq_bvh_export writes 2-D boxes (4 doubles each);
 * axsq_bvh_bounds reports z = 0. */
int axsq_bvh_build2d(axsq_bvh** out, const double* boxes, int64_t n, double scale,
                     int fp_mode, void* stream);
/* 2 or 3: the dimension the handle was built in (0 for NULL). */
int axsq_bvh_dimension(const axsq_bvh* bvh);

/* Builds a float tree over n boxes, AoS of 2 * ndims floats (min, max), ndims
 * 2 or 3. Replaces spin::BVH<NDIMS, Exec, float>::initialize (BVH.hpp:129-130
 * FloatType, :426-477; the reference's tests instantiate every scenario with
 * float, spin/tests/spin_bvh.cpp:1560-1669): the scaling about the centroid,
 * the bounds, the Morton normalisation and the ray test run in float, and
 * the scale factor is a float (BVH::setScaleFactor(FloatType), default
 * 1.000123f). Boxes are kept as their float values; axsq_bvh_bounds,
 * axsq_bvh_export and axsq_bvh_device_arrays report them as doubles (exact).
 * Queries go through the axsq_bvh_find*_f32 entry points, with float
 * coordinates and a float tolerance (FLT_EPSILON by default, BVH.hpp:411-412);
 * the double ones refuse a float tree (AXSQ_ERR_STATE), and the other way
 * round. */
int axsq_bvh_build_f32(axsq_bvh** out, const float* boxes, int64_t n, int ndims, float scale,
                       void* stream);
/* axsq_bvh_rebuild for a float tree (the dimension it was built in). */
int axsq_bvh_rebuild_f32(axsq_bvh* bvh, const float* boxes, int64_t n, float scale,
                         void* stream);
/* 32 for a float tree, 64 for a double one, 0 for NULL. */
int axsq_bvh_precision(const axsq_bvh* bvh);

/* One inner node of the tree on the device (128 bytes). Child link >= 0: an
 * inner node index; < 0: -(leaf position + 1). flags bit 0 / bit 1: the left /
 * right child box is valid (BoundingBox::isValid). Boxes are the scaled boxes
 * the BVH stores, min xyz then max xyz (a 2-D tree holds z = [0, 0]). [lo, hi]
 * is the range of leaf positions under the node, split after mid. */
typedef struct axsq_node {
  double l[6];
  double r[6];
  int32_t lc, rc;
  int32_t flags;
  int32_t lo, hi, mid;
  int32_t pad[2];
} axsq_node;

/* The device arrays a device-side walk reads: BVH::getTraverser
 * (src/axom/spin/BVH.hpp:305-319) -> LinearBVHTraverser
 * (src/axom/spin/policy/LinearBVH.hpp:57-109). *nodes: the ninner inner
 * nodes, root 0; *leaf_ids: the original box id of each of the ninner + 1
 * leaf positions. They stay valid until the tree is rebuilt or destroyed. */
int axsq_bvh_device_arrays(const axsq_bvh* bvh, const axsq_node** nodes,
                           const int32_t** leaf_ids, int32_t* ninner);
/* Builds again over new boxes into an existing tree (3-D or 2-D, as it was
 * built), keeping its device arrays when the box count is unchanged: calling
 * spin::BVH::initialize again on an initialized BVH (src/axom/spin/BVH.hpp:
 * 426-477 replaces the tree in place). Must be called on the device the tree
 * was built on. After an error the tree may only be destroyed. */
int axsq_bvh_rebuild(axsq_bvh* bvh, const double* boxes, int64_t n, double scale, int fp_mode,
                     void* stream);

/* Sets the memory of the candidate arrays axsq_bvh_find returns (one of
 * AXSQ_ALLOC_*): BVH::setAllocatorID (BVH.hpp:251-261). */
int axsq_bvh_set_allocator(axsq_bvh* bvh, int allocator_id);
int axsq_bvh_get_allocator(const axsq_bvh* bvh);

/* Writes the tree as a legacy-VTK unstructured grid of boxes (hexahedra, or
 * quads for a 2-D tree): the root box at level 0, then every inner node's two
 * child boxes at their depth, depth first, left before right, with a "level"
 * cell scalar. BVH::writeVtkFile -> LinearBVH::writeVtkFileImpl
 * (src/axom/spin/BVH.hpp:405,572-577; policy/LinearBVH.hpp:404-462;
 * internal/linear_bvh/bvh_vtkio.hpp), numbers written as a C++ stream writes
 * them by default (6 significant digits). Host file output. */
int axsq_bvh_write_vtk(const axsq_bvh* bvh, const char* file);

/* Root box of the BVH: BVH::getBounds (BVH.hpp:297-308). Host outputs. */
int axsq_bvh_bounds(const axsq_bvh* bvh, double lo[3], double hi[3]);
/* Number of leaves after padding. */
int64_t axsq_bvh_num_leaves(const axsq_bvh* bvh);

/* Copies the tree out in the reference's own layout (device outputs, any may
 * be NULL): inner_nodes 2(n-1) boxes, children 2(n-1) int32 (2*child for an
 * inner child, -(pos+1) for a leaf), leaf_nodes n int32, mcodes n uint32
 * (LinearBVH.hpp:170-187, RadixTree.hpp:28-63). */
int axsq_bvh_export(const axsq_bvh* bvh, double* inner_nodes, int32_t* children,
                    int32_t* leaf_nodes, uint32_t* mcodes, void* stream);

/* Candidate search, two-phase form with caller-owned candidate storage:
 * the count pass, exclusive scan and fill pass of the reference's GPU path,
 * LinearBVH::findCandidatesImpl (LinearBVH.hpp:273-402, count :296-330, fill
 * :332-367) behind BVH::findPoints / findBoundingBoxes / findRays
 * (BVH.hpp:482-568).
 *   kind AXSQ_QUERY_POINTS: prims = x pointer, y/z/stride as in the header notes
 *   kind AXSQ_QUERY_BOXES : prims = AoS boxes (6 doubles each), y=z=NULL, stride 6
 *   kind AXSQ_QUERY_RAYS  : prims = AoS rays {origin xyz, unit direction xyz},
 *                           y=z=NULL, stride 6; tol = BVH::setTolerance value
 * Phase 1 (axsq_bvh_find_count) writes counts[nq], offsets[nq] (exclusive
 * scan) and *total. Phase 2 (axsq_bvh_find_fill) searches again and writes
 * candidates: for query i, ids at [offsets[i], offsets[i]+counts[i]) in
 * traversal order, never more than counts[i] of them. When its search finds
 * another number of candidates for some query than counts[i] (the queries,
 * the tolerance or the tree changed after phase 1), it returns
 * AXSQ_ERR_INVALID_ARG; no id is written outside the queries' ranges. */
int axsq_bvh_find_count(const axsq_bvh* bvh, int kind, const double* x, const double* y,
                        const double* z, int64_t stride, int64_t nq, double tol,
                        int32_t* offsets, int32_t* counts, int64_t* total, void* stream);
int axsq_bvh_find_fill(const axsq_bvh* bvh, int kind, const double* x, const double* y,
                       const double* z, int64_t stride, int64_t nq, double tol,
                       const int32_t* offsets, const int32_t* counts, int32_t* candidates,
                       void* stream);

/* One call into caller-owned storage: counts, offsets and *total as phase 1,
 * then, when *total <= capacity, the candidates as phase 2 writes them. The
 * search runs once: the count pass keeps each query's first 64 ids (fewer
 * when that would pass 512 MB) in the calling thread's workspace and only
 * queries with more are searched again. When *total > capacity nothing is
 * written to candidates (call axsq_bvh_find_fill with room for *total ids, or
 * this function again). No reference counterpart as one call: the count and
 * fill of LinearBVH::findCandidatesImpl (LinearBVH.hpp:296-367) in one. */
int axsq_bvh_find_into(const axsq_bvh* bvh, int kind, const double* x, const double* y,
                       const double* z, int64_t stride, int64_t nq, double tol, int32_t* offsets,
                       int32_t* counts, int32_t* candidates, int64_t capacity, int64_t* total,
                       void* stream);

/* axsq_bvh_find_into without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, results and errors as axsq_bvh_find_into, for at most 1,048,576
 * queries (more: AXSQ_ERR_INVALID_ARG, nothing written). Its scratch (kept
 * ids, counts, the scan's sums and totals) is a fixed 32 MiB array in device
 * memory, one per GPU, held until the process ends; it keeps up to 64 ids per
 * query from the count pass (fewer above 127,084 queries, at least 5).
 * Every call on one GPU uses that GPU's one array: calls that overlap on the
 * same GPU (other threads or streams) race and must be serialised by the
 * caller. */
int axsq_bvh_find_into_noalloc(const axsq_bvh* bvh, int kind, const double* x, const double* y,
                               const double* z, int64_t stride, int64_t nq, double tol,
                               int32_t* offsets, int32_t* counts, int32_t* candidates,
                               int64_t capacity, int64_t* total, void* stream);

/* The four searches above for a float tree (spin::BVH<NDIMS, Exec, float>::
 * findPoints / findBoundingBoxes / findRays, BVH.hpp:482-568): the same
 * arguments with float queries (strides in floats) and a float tolerance. */
int axsq_bvh_find_count_f32(const axsq_bvh* bvh, int kind, const float* x, const float* y,
                            const float* z, int64_t stride, int64_t nq, float tol,
                            int32_t* offsets, int32_t* counts, int64_t* total, void* stream);
int axsq_bvh_find_fill_f32(const axsq_bvh* bvh, int kind, const float* x, const float* y,
                           const float* z, int64_t stride, int64_t nq, float tol,
                           const int32_t* offsets, const int32_t* counts, int32_t* candidates,
                           void* stream);
int axsq_bvh_find_into_f32(const axsq_bvh* bvh, int kind, const float* x, const float* y,
                           const float* z, int64_t stride, int64_t nq, float tol,
                           int32_t* offsets, int32_t* counts, int32_t* candidates,
                           int64_t capacity, int64_t* total, void* stream);
int axsq_bvh_find_f32(const axsq_bvh* bvh, int kind, const float* x, const float* y,
                      const float* z, int64_t stride, int64_t nq, float tol, int32_t* offsets,
                      int32_t* counts, int32_t** candidates, int64_t* total, void* stream);

/* One-call form matching BVH::find* semantics: counts and offsets as above;
 * the library allocates *candidates (*total entries, in the memory set by
 * axsq_bvh_set_allocator, device memory by default; release with axsq_free),
 * as the reference's BVH allocates the candidate array (LinearBVH.hpp:
 * 338-341). It searches once, as axsq_bvh_find_into does. */
int axsq_bvh_find(const axsq_bvh* bvh, int kind, const double* x, const double* y,
                  const double* z, int64_t stride, int64_t nq, double tol, int32_t* offsets,
                  int32_t* counts, int32_t** candidates, int64_t* total, void* stream);

/* ---- SignedDistance -----------------------------------------------------
 * Replaces quest::SignedDistance<3,Exec>(mesh, isWatertight, computeSign)
 * (src/axom/quest/SignedDistance.hpp:411-504): mesh bounds, per-triangle
 * boxes and a BVH with the default scale 1.000123. The handle COPIES the
 * mesh (the reference keeps a pointer to the mint mesh instead).
 * x,y,z: device SoA vertex coordinates (nverts); tris: device int32
 * connectivity (3*ntris), as mint::UnstructuredMesh<SINGLE_SHAPE> holds them
 * (mint/mesh/UnstructuredMesh.hpp:775-783,1256). */
int axsq_sdf_create(axsq_sdf** out, const double* x, const double* y, const double* z,
                    int64_t nverts, const int32_t* tris, int64_t ntris, int watertight,
                    int compute_sign, int fp_mode, void* stream);
/* The same for a surface mesh of linear triangle and quad cells (a
 * mint::UnstructuredMesh of TRIANGLE or QUAD, SINGLE_SHAPE or MIXED_SHAPE;
 * SignedDistance.hpp:609-660, SignedDistance.cpp:16-45). conn: device int32
 * node ids; cell_offsets: ncells + 1 int32 offsets into conn (host or device),
 * each cell 3 or 4 nodes. A quad (n0, n1, n2, n3) is checked as the
 * triangles (n0, n1, n2) then (n0, n2, n3). A mesh of triangles only takes
 * the axsq_sdf_create path; one with quads walks the tree in one pass on FP64
 * nodes (fp_mode is then ignored) and has no adjoint (axsq_sdf_backward
 * returns AXSQ_ERR_INVALID_ARG). */
int axsq_sdf_create_cells(axsq_sdf** out, const double* x, const double* y, const double* z,
                          int64_t nverts, const int32_t* conn, const int32_t* cell_offsets,
                          int64_t ncells, int watertight, int compute_sign, int fp_mode,
                          void* stream);
/* The same from an STL file read as quest::STLReader reads it (see
 * axsq_stl_read): one vertex per facet corner, implicit connectivity. This is
 * signed_distance_init(file) without the singleton
 * (quest/interface/signed_distance.cpp:109-160 ->
 * interface/internal/QuestHelpers.cpp:287-325). AXSQ_ERR_INVALID_ARG when the
 * file cannot be read or holds no triangle. */
int axsq_sdf_create_from_stl(axsq_sdf** out, const char* file, int watertight,
                             int compute_sign, int fp_mode, void* stream);
int axsq_sdf_destroy(axsq_sdf* sdf);
/* Mesh bounding box (signed_distance_get_mesh_bounds,
 * quest/interface/signed_distance.cpp:468-488). Host outputs. */
int axsq_sdf_mesh_bounds(const axsq_sdf* sdf, double lo[3], double hi[3]);
/* Replaces the vertex coordinates (same connectivity) and rebuilds the BVH:
 * the rebuild step of a shape-fitting loop. No reference counterpart beyond
 * SignedDistance::setMesh (SignedDistance.hpp:428-504). */
int axsq_sdf_update_vertices(axsq_sdf* sdf, const double* x, const double* y,
                             const double* z, void* stream);

/* Signed distances for nq query points.
 * Replaces SignedDistance::computeDistances (SignedDistance.hpp:529-605) and,
 * through it, signed_distance_evaluate (quest/interface/signed_distance.cpp:432-465).
 * fp_mode: AXSQ_FP64 (plain) or AXSQ_FP32_TRAVERSAL (needs a handle created
 * with AXSQ_FP32_TRAVERSAL). Outputs (device, NULL to skip):
 *   dist[nq]        signed distance (sign carried also for zero distances)
 *   cp[3*nq]        closest surface point (AoS)
 *   normal[3*nq]    unit (pseudo-)normal at the closest point (AoS)
 *   elem[nq]        winning triangle id (input of axsq_sdf_backward); -1 when
 *                   no triangle was taken (a NaN or infinite coordinate,
 *                   whose dist is then -sqrt(DBL_MAX) or +sqrt(DBL_MAX))    */
int axsq_sdf_eval(const axsq_sdf* sdf, const double* x, const double* y, const double* z,
                  int64_t stride, int64_t nq, int fp_mode, double* dist, double* cp,
                  double* normal, int32_t* elem, void* stream);

/* One query point, host in and out: SignedDistance::computeDistance(pt)
 * and computeDistance(pt, closestPnt, surfaceNormal)
 * (src/axom/quest/SignedDistance.hpp:508-524), through the batched kernel.
 * cp and normal may be NULL. Uses a small device buffer of the calling
 * thread, so it allocates nothing after its first call on a device. */
int axsq_sdf_eval_point(const axsq_sdf* sdf, const double q[3], int fp_mode, double* dist,
                        double cp[3], double normal[3]);

/* Adjoint of axsq_sdf_eval (no reference counterpart):
 *   grad_q[3*nq]     = g_i * sgn_i * (q_i - cp_i)/|q_i - cp_i|        (AoS)
 *   grad_v[3*nverts] = - sum over queries of w_k * that, w = barycentric
 *                      weights of cp_i on triangle elem_i              (AoS)
 * gout[nq] is the upstream gradient (float). Sums are deterministic: per
 * triangle and corner in ascending query order, then per vertex in
 * ascending triangle order, so repeated calls are bitwise identical.
 * A query with elem_i = -1 (no closest triangle) gets grad_q_i = 0 and adds
 * nothing to grad_v; any other elem_i outside [0, ntris) fails the call with
 * AXSQ_ERR_INVALID_ARG. Either output may be NULL. */
int axsq_sdf_backward(const axsq_sdf* sdf, const double* x, const double* y, const double* z,
                      int64_t stride, int64_t nq, const double* dist, const int32_t* elem,
                      const float* gout, double* grad_q, double* grad_v, void* stream);

/* ---- STL reader ---------------------------------------------------------
 * quest::STLReader::read (src/axom/quest/readers/STLReader.cpp:40-205): a file
 * of at least 84 bytes whose size is 84 + 50 * (the int32 at byte 80) is
 * binary, whose float coordinates are widened to double; any other is ASCII,
 * where each token "vertex" is followed by three coordinates read as double.
 * *nodes receives 3 * *nnodes doubles (x, y, z of each facet corner, AoS, host
 * memory; release with axsq_host_free); *nfaces = *nnodes / 3, with the
 * implicit connectivity (3i, 3i+1, 3i+2) of STLReader::getMesh (:207-254).
 * AXSQ_ERR_INVALID_ARG when the file cannot be opened. */
int axsq_stl_read(const char* file, double** nodes, int64_t* nnodes, int64_t* nfaces);

/* ---- measurement --------------------------------------------------------
 * When enabled, the library records HIP events around its main kernels on
 * the stream they run on. axsq_prof_get returns, for kernel `name`, the
 * number of launches and their summed duration in milliseconds. */
int axsq_prof_enable(int on);
int axsq_prof_reset(void);
int axsq_prof_get(const char* name, int64_t* launches, double* total_ms);

/* Traversal statistics of the candidate search when enabled (no reference
 * counterpart): the inner-node visits of the last count pass (axsq_bvh_find,
 * axsq_bvh_find_into, axsq_bvh_find_count) summed over its queries. Every
 * walk follows the reference's order, so this is the reference traversal's
 * count. */
int axsq_bvh_stats_enable(axsq_bvh* bvh, int on);
int axsq_bvh_stats_get(const axsq_bvh* bvh, int64_t* visits);

/* Per-query traversal statistics of the last FP64 axsq_sdf_eval when
 * enabled: summed inner-node visits and leaf (triangle) checks. */
int axsq_sdf_stats_enable(axsq_sdf* sdf, int on);
int axsq_sdf_stats_get(const axsq_sdf* sdf, int64_t* visits, int64_t* leaves);

/* Number of query walks of the last axsq_sdf_eval that ran past the first
 * pass's step budget and were finished by the second pass (no reference
 * counterpart: the reference runs every walk in one pass). */
int axsq_sdf_last_suspended(const axsq_sdf* sdf, int64_t* count);

/* Number of query walks the last axsq_sdf_eval ran from the start beside its
 * first pass: the longest walks of the previous call on the same query
 * arrays (same x, y, z pointers, stride and count). Scheduling only; every
 * query is walked once either way (no reference counterpart). */
int axsq_sdf_last_early(const axsq_sdf* sdf, int64_t* count);

/* ---- Triangle intersection and mesh self-intersection ----------------------
 * primal::intersect(Triangle3 t1, Triangle3 t2, includeBoundary, EPS)
 * (src/axom/primal/operators/intersect.hpp:65-71 ->
 * primal/operators/detail/intersect_impl.hpp:155-382) over n pairs: t1, t2
 * device AoS corners (9 doubles per triangle: A, B, C); out device uint8
 * (1: the pair intersects). */
int axsq_tri_tri_intersect(const double* t1, const double* t2, int64_t n, int include_boundary,
                           double eps, uint8_t* out, void* stream);

/* quest::findTriMeshIntersectionsBVH<Exec, double>(mesh, intersections,
 * degenerateIndices, intersectionThreshold) (src/axom/quest/MeshTester.hpp:
 * 66-87 -> quest/detail/MeshTester_detail.hpp:159-341). x, y, z: device SoA
 * vertices; tris: device int32 connectivity (3*ntris). The library allocates
 * *first and *second (*npairs entries each) and *degenerate (*ndegen
 * entries), all device int32; release each with axsq_free. Pairs (first[k],
 * second[k]) have first < second and come in the reference's sequential
 * order: by first, then by second's place in first's BVH candidate list.
 * Degenerate triangles (area within 1e-12 of zero) are listed ascending and,
 * as in the reference's BVH tester, still take part in the pair tests. */
int axsq_mesh_self_intersections(const double* x, const double* y, const double* z,
                                 int64_t nverts, const int32_t* tris, int64_t ntris,
                                 double threshold, int32_t** first, int32_t** second,
                                 int64_t* npairs, int32_t** degenerate, int64_t* ndegen,
                                 void* stream);

/* Candidate pairs between two meshes of cells with nodes_per_cell nodes each
 * (hexahedra: 8): findCandidatesBVH of src/axom/quest/examples/
 * quest_candidates_example.cpp:397-495. Each cell's bounding box
 * (primal::compute_bounding_box: the box of its nodes), a BVH over the insert
 * mesh's boxes with the default scale (1.000123), findBoundingBoxes for every
 * query box, and the pair list: first[k] the query cell, second[k] an insert
 * cell whose scaled box overlaps it, grouped by query cell ascending and, for
 * one query cell, in BVH traversal order (the order of the reference's
 * candidate array). Vertices are device SoA, cells device int32
 * (nodes_per_cell per cell, 1..8). The library allocates *first and *second
 * (*npairs entries each, device int32); release each with axsq_free. */
int axsq_mesh_candidates(const double* ix, const double* iy, const double* iz, int64_t inverts,
                         const int32_t* icells, int64_t nicells, const double* qx,
                         const double* qy, const double* qz, int64_t qnverts,
                         const int32_t* qcells, int64_t nqcells, int nodes_per_cell,
                         int32_t** first, int32_t** second, int64_t* npairs, void* stream);

#ifdef __cplusplus
}
#endif
#endif
