// axom_mi355x.hpp -- C++ host interface of the MI355X spatial-query path.
//
// Mirrors the reference's user-facing C++ API for this path over the C-ABI in
// axsq.h, so user code keeps its shape:
//   axom::spin::BVH<3, axom::MI355X_EXEC>          (src/axom/spin/BVH.hpp:129-419)
//   axom::quest::SignedDistance<3, axom::MI355X_EXEC>
//                                                 (src/axom/quest/SignedDistance.hpp:147-397)
//   axom::execution_space<axom::MI355X_EXEC>       (src/axom/core/execution/execution_space.hpp:77-103)
// Bulk data are device pointers wrapped in axom::ArrayView-like views, as the
// reference requires for GPU exec spaces (BVH.hpp:198-203). Candidate arrays
// are allocated by the BVH (LinearBVH.hpp:338-341) in an axom::Array-like
// owning device array. Misuse throws axom::mi355x_error where the reference
// aborts via SLIC_ERROR.
//
// Header-only; link with -laxsq (axom_amd/lib/libaxsq.so) and the HIP runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "axsq.h"
#include "axsq_quest.h"

namespace axom {

using IndexType = std::int32_t;  // core/Types.hpp:63-65 (default, 32-bit)

class mi355x_error : public std::runtime_error {
 public:
  mi355x_error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

namespace detail {
inline void check(int rc) {
  if (rc != AXSQ_OK) throw mi355x_error(rc, axsq_last_error());
}
inline void hip_check(hipError_t e) {
  if (e != hipSuccess) throw mi355x_error(AXSQ_ERR_HIP, hipGetErrorString(e));
}
// the quest C interface reports misuse through axsq_last_error() (empty after
// a call that succeeded)
inline void quest_check() {
  const char* e = axsq_last_error();
  if (e && e[0]) throw mi355x_error(AXSQ_ERR_STATE, e);
}
}  // namespace detail

// The single MI355X execution space.
struct MI355X_EXEC {};

template <typename ExecSpace>
struct execution_space;

template <>
struct execution_space<MI355X_EXEC> {
  static constexpr bool valid() noexcept { return true; }
  static constexpr bool async() noexcept { return false; }  // calls return when done
  static constexpr bool onDevice() noexcept { return true; }
  static constexpr char* name() noexcept { return (char*)"[MI355X_EXEC]"; }
  // allocator ids: AXSQ_ALLOC_DEVICE (0, the default: hipMalloc),
  // AXSQ_ALLOC_MANAGED (1) and AXSQ_ALLOC_PINNED_HOST (2); all are memory the
  // device reads and writes
  static int allocatorID() noexcept { return AXSQ_ALLOC_DEVICE; }
  static bool usesAllocId(int id) noexcept {
    return id == AXSQ_ALLOC_DEVICE || id == AXSQ_ALLOC_MANAGED || id == AXSQ_ALLOC_PINNED_HOST;
  }
};

// Non-owning view of device memory (axom::ArrayView on device memory).
template <typename T>
class ArrayView {
 public:
  ArrayView() = default;
  ArrayView(T* data, IndexType size) : data_(data), size_(size) {}
  T* data() const { return data_; }
  IndexType size() const { return size_; }

 private:
  T* data_ = nullptr;
  IndexType size_ = 0;
};

// Owning device array (axom::Array with a device allocator).
template <typename T>
class Array {
 public:
  Array() = default;
  explicit Array(IndexType n) { resize(n); }
  Array(const Array&) = delete;
  Array& operator=(const Array&) = delete;
  Array(Array&& o) noexcept { swap(o); }
  Array& operator=(Array&& o) noexcept {
    swap(o);
    return *this;
  }
  ~Array() { reset(); }
  void resize(IndexType n) {
    reset();
    if (n > 0) detail::hip_check(hipMalloc((void**)&data_, sizeof(T) * (size_t)n));
    size_ = n;
  }
  // Takes ownership of library-allocated memory (released with axsq_free).
  void adopt(T* p, IndexType n) {
    reset();
    data_ = p;
    size_ = n;
    lib_owned_ = true;
  }
  T* data() const { return data_; }
  IndexType size() const { return size_; }
  ArrayView<T> view() const { return ArrayView<T>(data_, size_); }

 private:
  void reset() {
    if (data_) {
      if (lib_owned_)
        (void)axsq_free(data_);
      else
        (void)hipFree(data_);
    }
    data_ = nullptr;
    size_ = 0;
    lib_owned_ = false;
  }
  void swap(Array& o) {
    std::swap(data_, o.data_);
    std::swap(size_, o.size_);
    std::swap(lib_owned_, o.lib_owned_);
  }
  T* data_ = nullptr;
  IndexType size_ = 0;
  bool lib_owned_ = false;
};

namespace primal {
// Device layouts of primal::BoundingBox<double,N> (min, max),
// primal::Point<double,N> and primal::Ray<double,N> (origin, unit direction);
// the library reads them as plain doubles.
struct BoundingBox3 {
  double min[3];
  double max[3];
};
struct Point3 {
  double x[3];
};
struct Vector3 {  // primal::Vector<double,3>
  double x[3];
};
struct Ray3 {
  double origin[3];
  double dir[3];
};
struct BoundingBox2 {
  double min[2];
  double max[2];
};
struct Point2 {
  double x[2];
};
struct Ray2 {
  double origin[2];
  double dir[2];
};
// primal::Ray stores a unit direction (Ray.hpp:122-128); make_ray normalises
// the way Vector::unitVector does (Vector.hpp:477-490).
template <int N, class RayT>
inline RayT make_ray_n(const double* o, const double* d) {
  RayT r;
  double len_sq = 0.0;
  for (int k = 0; k < N; ++k) {
    r.origin[k] = o[k];
    len_sq += d[k] * d[k];
    r.dir[k] = (k == 0) ? 1.0 : 0.0;
  }
  if (len_sq >= 1e-50) {
    const double inv = 1. / std::sqrt(len_sq);
    for (int k = 0; k < N; ++k) r.dir[k] = d[k] * inv;
  }
  return r;
}
inline Ray3 make_ray(const double o[3], const double d[3]) { return make_ray_n<3, Ray3>(o, d); }
inline Ray2 make_ray2(const double o[2], const double d[2]) { return make_ray_n<2, Ray2>(o, d); }

// The float layouts (primal::BoundingBox<float,N>, Point<float,N>,
// Ray<float,N>) of spin::BVH<NDIMS, MI355X_EXEC, float>.
struct BoundingBox3f {
  float min[3];
  float max[3];
};
struct Point3f {
  float x[3];
};
struct Ray3f {
  float origin[3];
  float dir[3];
};
struct BoundingBox2f {
  float min[2];
  float max[2];
};
struct Point2f {
  float x[2];
};
struct Ray2f {
  float origin[2];
  float dir[2];
};
// Ray<float>'s unit direction (Vector<float>::unitVector, Vector.hpp:477-490,
// NumericArray.hpp:498-514): the squared norm summed in float, then each
// component times 1 / double(float(sqrt)), rounded to float.
template <int N, class RayT>
inline RayT make_ray_nf(const float* o, const float* d) {
  RayT r;
  float len_sq = 0.0f;
  for (int k = 0; k < N; ++k) {
    r.origin[k] = o[k];
    len_sq += d[k] * d[k];
    r.dir[k] = (k == 0) ? 1.0f : 0.0f;
  }
  if ((double)len_sq >= 1e-50) {
    const float s = (float)std::sqrt((double)len_sq);
    const double inv = 1. / (double)s;
    for (int k = 0; k < N; ++k) r.dir[k] = (float)((double)d[k] * inv);
  }
  return r;
}
inline Ray3f make_ray(const float o[3], const float d[3]) { return make_ray_nf<3, Ray3f>(o, d); }
inline Ray2f make_ray2(const float o[2], const float d[2]) { return make_ray_nf<2, Ray2f>(o, d); }
}  // namespace primal

namespace spin {

enum BVHReturnCodes { BVH_BUILD_FAILED = -1, BVH_BUILD_OK = 0 };  // BVH.hpp:39-43

template <int NDIMS, typename ExecSpace, typename FloatType = double>
class BVH;

namespace detail {
template <int NDIMS, typename T = double>
struct BVHTypes;
template <>
struct BVHTypes<3, double> {
  using BoxType = primal::BoundingBox3;
  using PointType = primal::Point3;
  using RayType = primal::Ray3;
};
template <>
struct BVHTypes<2, double> {
  using BoxType = primal::BoundingBox2;
  using PointType = primal::Point2;
  using RayType = primal::Ray2;
};
template <>
struct BVHTypes<3, float> {
  using BoxType = primal::BoundingBox3f;
  using PointType = primal::Point3f;
  using RayType = primal::Ray3f;
};
template <>
struct BVHTypes<2, float> {
  using BoxType = primal::BoundingBox2f;
  using PointType = primal::Point2f;
  using RayType = primal::Ray2f;
};

// The device-side walk of spin::BVH::getTraverser (LinearBVHTraverser,
// policy/LinearBVH.hpp:57-109, over bvh_traverse, internal/linear_bvh/
// bvh_traverse.hpp:66-154): for HIP kernels of the caller. The walk keeps a
// 64-entry stack; at each inner node it tests both valid child boxes with the
// predicate, goes down the one that passes (both: the preferred one first,
// the other pushed), holds the first leaf it reaches until a second leaf or
// the end of a descent, and then calls the leaf action on the held leaves in
// order: lf(leaf position, leaf id array), the id array mapping a leaf
// position to the original box index. The overload for the BVH's point type
// prefers the child whose box centre is nearer (an invalid right box never
// wins); every other primitive keeps the left-first order. Device code only:
// the arrays are device memory, valid until the tree is rebuilt or destroyed.
// traverse_tree returns false when a push would not fit in the stack (STACK
// entries, the first the end marker: 63 pending nodes by default): the walk
// stops there, so the leaf actions it made are a prefix of the full walk's
// (the reference has no check; its 64-entry stack would overflow).
// T: the tree's FloatType; a float tree's boxes reach the predicate as float
// boxes (its node arrays hold float values) and the nearer-first test runs
// in float, as LinearBVHTraverser<float, NDIMS> computes it.
template <int NDIMS, int STACK = 64, typename T = double>
class MI355XTraverser {
  static_assert(STACK >= 2, "the walk's stack needs room for its end marker and a node");

 public:
  using BoxType = typename BVHTypes<NDIMS, T>::BoxType;
  using PointType = typename BVHTypes<NDIMS, T>::PointType;

  MI355XTraverser() = default;
  MI355XTraverser(const axsq_node* nodes, const std::int32_t* leaf_ids, std::int32_t ninner)
      : m_nodes(nodes), m_leaf_ids(leaf_ids), m_ninner(ninner) {}

  template <typename LeafAction, typename Predicate>
  __device__ bool traverse_tree(const PointType& p, LeafAction&& lf, Predicate&& predicate) const {
    return walk(p, lf, predicate, [&p](const BoxType& l, const BoxType& r, bool r_valid) {
      if (!r_valid) return false;
      T dl = 0, dr = 0;
      for (int d = 0; d < NDIMS; ++d) {
        // centroid T(0.5 * (min + max)), the sum in T (Point.hpp:279-290)
        const T cl = static_cast<T>(0.5 * static_cast<T>(l.min[d] + l.max[d]));
        const T cr = static_cast<T>(0.5 * static_cast<T>(r.min[d] + r.max[d]));
        const T el = p.x[d] - cl;
        const T er = p.x[d] - cr;
        dl += el * el;
        dr += er * er;
      }
      return dl > dr;
    });
  }

  template <typename Primitive, typename LeafAction, typename Predicate>
  __device__ bool traverse_tree(const Primitive& p, LeafAction&& lf, Predicate&& predicate) const {
    return walk(p, lf, predicate, [](const BoxType&, const BoxType&, bool) { return false; });
  }

 private:
  __device__ static BoxType box_of(const double* b) {
    BoxType out;
    for (int d = 0; d < NDIMS; ++d) {
      out.min[d] = static_cast<T>(b[d]);
      out.max[d] = static_cast<T>(b[3 + d]);
    }
    return out;
  }

  template <typename Primitive, typename LeafAction, typename Predicate, typename Pref>
  __device__ bool walk(const Primitive& p, LeafAction& lf, Predicate& predicate, Pref&& pref) const {
    constexpr std::int32_t kDepth = STACK;
    constexpr std::int32_t kEnd = -2000000000;  // below every leaf link
    std::int32_t stack[kDepth];
    std::int32_t top = 0;
    stack[0] = kEnd;
    std::int32_t node = 0, held = 0;  // held < 0: a leaf waiting for its action
    const std::int32_t* ids = m_leaf_ids;
    while (node != kEnd) {
      while (node >= 0) {  // descend through inner nodes
        const axsq_node& nd = m_nodes[node];
        const BoxType lb = box_of(nd.l), rb = box_of(nd.r);
        const bool go_l = (nd.flags & 1) ? predicate(p, lb) : false;
        const bool go_r = (nd.flags & 2) ? predicate(p, rb) : false;
        if (!go_l && !go_r) {
          node = stack[top--];
        } else if (go_l && go_r) {
          const bool right_first = pref(lb, rb, (nd.flags & 2) != 0);
          node = right_first ? nd.rc : nd.lc;
          if (top + 1 >= kDepth) return false;  // deeper than the reference's stack
          stack[++top] = right_first ? nd.lc : nd.rc;
        } else {
          node = go_l ? nd.lc : nd.rc;
        }
        if (node < 0 && held >= 0) {  // hold the first leaf, keep descending
          held = node;
          if (node != kEnd) node = stack[top--];
        }
      }
      while (held < 0 && held != kEnd) {  // the held leaves, in order
        lf(static_cast<std::int32_t>(-held - 1), ids);
        held = node;
        if (node < 0 && node != kEnd) node = stack[top--];
      }
      held = 0;
    }
    return true;
  }

  const axsq_node* m_nodes = nullptr;
  const std::int32_t* m_leaf_ids = nullptr;
  std::int32_t m_ninner = 0;
};

// The body of spin::BVH<NDIMS, MI355X_EXEC, T> for NDIMS 3 and 2 and T double
// or float: build over device boxes, find* over device queries. The BVH
// copies the boxes; it keeps no pointer to them.
template <int NDIMS, typename T = double>
class MI355XLinearBVH {
  static constexpr bool F32 = std::is_same<T, float>::value;
  static_assert(F32 || std::is_same<T, double>::value, "FloatType is double or float");

 public:
  using BoxType = typename BVHTypes<NDIMS, T>::BoxType;
  using PointType = typename BVHTypes<NDIMS, T>::PointType;
  using RayType = typename BVHTypes<NDIMS, T>::RayType;

  static constexpr T DEFAULT_SCALE_FACTOR = static_cast<T>(1.000123);  // BVH.hpp:410
  static constexpr T DEFAULT_TOLERANCE = std::numeric_limits<T>::epsilon();

  MI355XLinearBVH() = default;
  MI355XLinearBVH(const MI355XLinearBVH&) = delete;
  MI355XLinearBVH& operator=(const MI355XLinearBVH&) = delete;
  ~MI355XLinearBVH() { (void)axsq_bvh_destroy(m_bvh); }

  // BVH.hpp:426-477
  int initialize(const BoxType* boxes, IndexType numItems) {
    if (F32) return initialize_f32(reinterpret_cast<const float*>(boxes), numItems);
    const double* b = reinterpret_cast<const double*>(boxes);
    if (m_bvh) {  // build again in place, keeping the device arrays
      const int rc = axsq_bvh_rebuild(m_bvh, b, numItems, m_scaleFactor, m_fpMode, nullptr);
      if (rc != AXSQ_OK) {
        (void)axsq_bvh_destroy(m_bvh);
        m_bvh = nullptr;
        ::axom::detail::check(rc);
      }
      return BVH_BUILD_OK;
    }
    if (NDIMS == 2)
      ::axom::detail::check(axsq_bvh_build2d(&m_bvh, b, numItems, m_scaleFactor, m_fpMode, nullptr));
    else
      ::axom::detail::check(axsq_bvh_build(&m_bvh, b, numItems, m_scaleFactor, m_fpMode, nullptr));
    ::axom::detail::check(axsq_bvh_set_allocator(m_bvh, m_allocatorID));
    return BVH_BUILD_OK;
  }
  int initialize(ArrayView<const BoxType> boxes, IndexType numItems) {
    return initialize(boxes.data(), numItems);
  }

 private:
  int initialize_f32(const float* b, IndexType numItems) {
    if (m_bvh) {
      const int rc = axsq_bvh_rebuild_f32(m_bvh, b, numItems, (float)m_scaleFactor, nullptr);
      if (rc != AXSQ_OK) {
        (void)axsq_bvh_destroy(m_bvh);
        m_bvh = nullptr;
        ::axom::detail::check(rc);
      }
      return BVH_BUILD_OK;
    }
    ::axom::detail::check(
        axsq_bvh_build_f32(&m_bvh, b, numItems, NDIMS, (float)m_scaleFactor, nullptr));
    ::axom::detail::check(axsq_bvh_set_allocator(m_bvh, m_allocatorID));
    return BVH_BUILD_OK;
  }

 public:
  bool isInitialized() const { return m_bvh != nullptr; }

  // BVH.hpp:251-261: the memory of the candidate arrays find* returns, one of
  // the ids execution_space<MI355X_EXEC>::usesAllocId accepts
  void setAllocatorID(int id) {
    if (!execution_space<MI355X_EXEC>::usesAllocId(id))
      throw mi355x_error(AXSQ_ERR_INVALID_ARG, "allocator id not usable by MI355X_EXEC");
    m_allocatorID = id;
    if (m_bvh) ::axom::detail::check(axsq_bvh_set_allocator(m_bvh, id));
  }
  int getAllocatorID() const { return m_allocatorID; }
  // BVH.hpp:405,572-577: the tree's boxes as a legacy-VTK file
  void writeVtkFile(const std::string& fileName) const {
    if (!m_bvh) throw mi355x_error(AXSQ_ERR_STATE, "BVH is not initialized");
    ::axom::detail::check(axsq_bvh_write_vtk(m_bvh, fileName.c_str()));
  }
  void setScaleFactor(T s) { m_scaleFactor = s; }
  T getScaleFactor() const { return m_scaleFactor; }
  void setTolerance(T eps) { m_tolerance = eps; }
  T getTolerance() const { return m_tolerance; }
  // Precision of the traversal of a double tree (MI355X addition): AXSQ_FP64
  // (default, the reference's form) or AXSQ_FP32_TRAVERSAL; takes effect at
  // initialize(). A float tree walks its own float nodes.
  void setTraversalPrecision(int fp_mode) { m_fpMode = fp_mode; }

  // BVH.hpp:297-308; an invalid box before initialize().
  BoxType getBounds() const {
    BoxType b;
    double lo[3], hi[3];
    if (!m_bvh) {
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::numeric_limits<T>::max();
        hi[k] = std::numeric_limits<T>::lowest();
      }
    } else {
      ::axom::detail::check(axsq_bvh_bounds(m_bvh, lo, hi));
    }
    for (int k = 0; k < NDIMS; ++k) {
      b.min[k] = static_cast<T>(lo[k]);  // a float tree's bounds are floats
      b.max[k] = static_cast<T>(hi[k]);
    }
    return b;
  }

  // BVH.hpp:482-505 (AoS points)
  void findPoints(ArrayView<IndexType> offsets, ArrayView<IndexType> counts,
                  Array<IndexType>& candidates, IndexType numPts, const PointType* points) const {
    const T* p = reinterpret_cast<const T*>(points);
    find(AXSQ_QUERY_POINTS, p, p + 1, NDIMS == 3 ? p + 2 : nullptr, NDIMS, offsets, counts,
         candidates, numPts);
  }
  // ZipPoint-style SoA points (z is ignored, and may be null, in 2-D)
  void findPoints(ArrayView<IndexType> offsets, ArrayView<IndexType> counts,
                  Array<IndexType>& candidates, IndexType numPts, const T* x, const T* y,
                  const T* z = nullptr) const {
    find(AXSQ_QUERY_POINTS, x, y, z, 1, offsets, counts, candidates, numPts);
  }
  // BVH.hpp:510-536
  void findRays(ArrayView<IndexType> offsets, ArrayView<IndexType> counts,
                Array<IndexType>& candidates, IndexType numRays, const RayType* rays) const {
    find(AXSQ_QUERY_RAYS, reinterpret_cast<const T*>(rays), nullptr, nullptr, 2 * NDIMS, offsets,
         counts, candidates, numRays);
  }
  // BVH.hpp:541-568
  void findBoundingBoxes(ArrayView<IndexType> offsets, ArrayView<IndexType> counts,
                         Array<IndexType>& candidates, IndexType numBoxes,
                         const BoxType* boxes) const {
    find(AXSQ_QUERY_BOXES, reinterpret_cast<const T*>(boxes), nullptr, nullptr, 2 * NDIMS,
         offsets, counts, candidates, numBoxes);
  }

  const axsq_bvh* handle() const { return m_bvh; }

  // BVH.hpp:305-319
  using TraverserType = MI355XTraverser<NDIMS, 64, T>;
  TraverserType getTraverser() const { return getTraverserWithStack<64>(); }
  // The same walker with a stack of STACK entries (an addition: the
  // reference's walk always has 64)
  template <int STACK>
  MI355XTraverser<NDIMS, STACK, T> getTraverserWithStack() const {
    if (!m_bvh) throw mi355x_error(AXSQ_ERR_STATE, "BVH is not initialized");
    const axsq_node* nodes = nullptr;
    const std::int32_t* ids = nullptr;
    std::int32_t ninner = 0;
    ::axom::detail::check(axsq_bvh_device_arrays(m_bvh, &nodes, &ids, &ninner));
    return MI355XTraverser<NDIMS, STACK, T>(nodes, ids, ninner);
  }

 private:
  void find(int kind, const T* x, const T* y, const T* z, int64_t stride,
            ArrayView<IndexType> offsets, ArrayView<IndexType> counts,
            Array<IndexType>& candidates, IndexType n) const {
    // LinearBVH.hpp:284-285
    if (offsets.size() != n) throw mi355x_error(AXSQ_ERR_INVALID_ARG, "offsets length != n");
    if (counts.size() != n) throw mi355x_error(AXSQ_ERR_INVALID_ARG, "counts length != n");
    if (!m_bvh) throw mi355x_error(AXSQ_ERR_STATE, "BVH is not initialized");
    int32_t* cand = nullptr;
    int64_t total = 0;
    if (F32)
      ::axom::detail::check(axsq_bvh_find_f32(
          m_bvh, kind, reinterpret_cast<const float*>(x), reinterpret_cast<const float*>(y),
          reinterpret_cast<const float*>(z), stride, n, (float)m_tolerance, offsets.data(),
          counts.data(), &cand, &total, nullptr));
    else
      ::axom::detail::check(axsq_bvh_find(
          m_bvh, kind, reinterpret_cast<const double*>(x), reinterpret_cast<const double*>(y),
          reinterpret_cast<const double*>(z), stride, n, (double)m_tolerance, offsets.data(),
          counts.data(), &cand, &total, nullptr));
    candidates.adopt(cand, (IndexType)total);
  }

 protected:
  axsq_bvh* m_bvh = nullptr;
  T m_tolerance = DEFAULT_TOLERANCE;
  T m_scaleFactor = DEFAULT_SCALE_FACTOR;
  int m_fpMode = AXSQ_FP64;
  int m_allocatorID = execution_space<MI355X_EXEC>::allocatorID();
};
}  // namespace detail

// spin::BVH<3 or 2, MI355X_EXEC, double> and spin::BVH<3 or 2, MI355X_EXEC,
// float> (BVH.hpp:129-130: FloatType is the box, point and ray coordinate
// type, and the type the build and the ray test compute in)
template <int NDIMS, typename FloatType>
class BVH<NDIMS, MI355X_EXEC, FloatType> : public detail::MI355XLinearBVH<NDIMS, FloatType> {
  static_assert(NDIMS == 2 || NDIMS == 3, "spin::BVH is 2-D or 3-D");
  using Base = detail::MI355XLinearBVH<NDIMS, FloatType>;

 public:
  using typename Base::BoxType;
  BVH() = default;
  BVH(const BoxType* boxes, IndexType numItems,
      int allocatorID = execution_space<MI355X_EXEC>::allocatorID(),
      FloatType tolerance = Base::DEFAULT_TOLERANCE,
      FloatType scaleFactor = Base::DEFAULT_SCALE_FACTOR) {
    this->setAllocatorID(allocatorID);
    this->m_tolerance = tolerance;
    this->m_scaleFactor = scaleFactor;
    this->initialize(boxes, numItems);
  }
};

}  // namespace spin

namespace quest {

template <int NDIMS, typename ExecSpace>
class SignedDistance;

// quest::SignedDistance<3, MI355X_EXEC>. The mesh is given as device SoA
// coordinates and int32 triangle connectivity (the arrays a
// mint::UnstructuredMesh<SINGLE_SHAPE> of TRIANGLEs holds); it is copied.
template <>
class SignedDistance<3, MI355X_EXEC> {
 public:
  using PointType = primal::Point3;
  using VectorType = primal::Vector3;

  SignedDistance(const double* x, const double* y, const double* z, IndexType nnodes,
                 const IndexType* connectivity, IndexType ncells, bool isWatertight,
                 bool computeSign = true, int fp_mode = AXSQ_FP64)
      : m_fpMode(fp_mode) {
    detail::check(axsq_sdf_create(&m_sdf, x, y, z, nnodes, connectivity, ncells,
                                  isWatertight ? 1 : 0, computeSign ? 1 : 0, fp_mode, nullptr));
  }
  SignedDistance(const SignedDistance&) = delete;
  SignedDistance& operator=(const SignedDistance&) = delete;
  ~SignedDistance() { (void)axsq_sdf_destroy(m_sdf); }

  // SignedDistance.hpp:529-605 with AoS points; outputs are device arrays
  // (outClosestPts / outNormals may be null).
  void computeDistances(int npts, const PointType* queryPts, double* outSgnDist,
                        PointType* outClosestPts = nullptr,
                        VectorType* outNormals = nullptr) const {
    const double* p = reinterpret_cast<const double*>(queryPts);
    detail::check(axsq_sdf_eval(m_sdf, p, p + 1, p + 2, 3, npts, m_fpMode, outSgnDist,
                                reinterpret_cast<double*>(outClosestPts),
                                reinterpret_cast<double*>(outNormals), nullptr, nullptr));
  }
  // ZipPoint-style SoA points
  void computeDistances(int npts, const double* x, const double* y, const double* z,
                        double* outSgnDist) const {
    detail::check(axsq_sdf_eval(m_sdf, x, y, z, 1, npts, m_fpMode, outSgnDist, nullptr, nullptr,
                                nullptr, nullptr));
  }
  // SignedDistance.hpp:508-512: one host point, through the same batched
  // kernel (the library stages it in a small device buffer of this thread).
  double computeDistance(const PointType& pt) const {
    double d = 0.0;
    detail::check(axsq_sdf_eval_point(m_sdf, pt.x, m_fpMode, &d, nullptr, nullptr));
    return d;
  }
  double computeDistance(double x, double y, double z) const {
    return computeDistance(PointType{{x, y, z}});
  }
  // SignedDistance.hpp:517-524: also the closest point and the unit
  // (pseudo-)normal there
  double computeDistance(const PointType& queryPnt, PointType& closestPnt,
                         VectorType& surfaceNormal) const {
    double d = 0.0;
    detail::check(axsq_sdf_eval_point(m_sdf, queryPnt.x, m_fpMode, &d, closestPnt.x,
                                      surfaceNormal.x));
    return d;
  }

  const axsq_sdf* handle() const { return m_sdf; }

 private:
  axsq_sdf* m_sdf = nullptr;
  int m_fpMode = AXSQ_FP64;
};

// quest::findTriMeshIntersectionsBVH<MI355X_EXEC, double> (MeshTester.hpp:
// 66-87): the mesh as device SoA coordinates and int32 connectivity; results
// on the host, as the reference returns them (std::vector of index pairs,
// std::vector of degenerate indices), in the reference's sequential order.
template <typename ExecSpace, typename FloatType>
void findTriMeshIntersectionsBVH(const double* x, const double* y, const double* z,
                                 IndexType nnodes, const IndexType* connectivity,
                                 IndexType ncells,
                                 std::vector<std::pair<int, int>>& intersections,
                                 std::vector<int>& degenerateIndices,
                                 double intersectionThreshold = 1E-8) {
  static_assert(std::is_same<ExecSpace, MI355X_EXEC>::value && std::is_same<FloatType, double>::value,
                "this build provides findTriMeshIntersectionsBVH<MI355X_EXEC, double>");
  int32_t *first = nullptr, *second = nullptr, *degen = nullptr;
  int64_t npairs = 0, ndegen = 0;
  detail::check(axsq_mesh_self_intersections(x, y, z, nnodes, connectivity, ncells,
                                             intersectionThreshold, &first, &second, &npairs,
                                             &degen, &ndegen, nullptr));
  Array<int32_t> f, g, d;  // released on every exit path
  f.adopt(first, npairs);
  g.adopt(second, npairs);
  d.adopt(degen, ndegen);
  std::vector<int32_t> hf((size_t)npairs), hg((size_t)npairs), hd((size_t)ndegen);
  if (npairs) {
    detail::hip_check(hipMemcpy(hf.data(), first, sizeof(int32_t) * npairs, hipMemcpyDeviceToHost));
    detail::hip_check(hipMemcpy(hg.data(), second, sizeof(int32_t) * npairs, hipMemcpyDeviceToHost));
  }
  if (ndegen)
    detail::hip_check(hipMemcpy(hd.data(), degen, sizeof(int32_t) * ndegen, hipMemcpyDeviceToHost));
  intersections.resize((size_t)npairs);
  for (int64_t k = 0; k < npairs; ++k) intersections[(size_t)k] = {hf[(size_t)k], hg[(size_t)k]};
  degenerateIndices.assign(hd.begin(), hd.end());
}

// quest::STLReader (src/axom/quest/readers/STLReader.hpp): setFileName,
// read (0 or -1), the node and face counts, and the mesh as SoA coordinates
// plus the implicit connectivity of getMesh (STLReader.cpp:207-254). Binary
// coordinates are widened from float, ASCII ones read as double.
class STLReader {
 public:
  STLReader() = default;
  explicit STLReader(const std::string& fileName) : m_fileName(fileName) {}
  void setFileName(const std::string& fileName) { m_fileName = fileName; }
  const std::string& getFileName() const { return m_fileName; }
  int read() {
    clear();
    double* nodes = nullptr;
    int64_t nn = 0, nf = 0;
    if (axsq_stl_read(m_fileName.c_str(), &nodes, &nn, &nf) != AXSQ_OK) return -1;
    m_nodes.assign(nodes, nodes + 3 * nn);
    (void)axsq_host_free(nodes);
    m_numNodes = (IndexType)nn;
    m_numFaces = (IndexType)nf;
    return 0;
  }
  void clear() {
    m_nodes.clear();
    m_numNodes = m_numFaces = 0;
  }
  IndexType getNumNodes() const { return m_numNodes; }
  IndexType getNumFaces() const { return m_numFaces; }
  // x, y, z of every node (getNumNodes entries each) and 3 * getNumFaces
  // connectivity entries, on the host
  void getMesh(std::vector<double>& x, std::vector<double>& y, std::vector<double>& z,
               std::vector<IndexType>& conn) const {
    x.resize(m_numNodes);
    y.resize(m_numNodes);
    z.resize(m_numNodes);
    for (IndexType i = 0; i < m_numNodes; ++i) {
      x[i] = m_nodes[3 * (size_t)i];
      y[i] = m_nodes[3 * (size_t)i + 1];
      z[i] = m_nodes[3 * (size_t)i + 2];
    }
    conn.resize(3 * (size_t)m_numFaces);
    for (size_t k = 0; k < conn.size(); ++k) conn[k] = (IndexType)k;
  }

 private:
  std::string m_fileName;
  std::vector<double> m_nodes;
  IndexType m_numNodes = 0, m_numFaces = 0;
};

// The signed-distance interface (src/axom/quest/interface/signed_distance.hpp:
// 88-319) over the library's process-wide query (axsq_quest.h). Misuse throws
// mi355x_error where the reference aborts through SLIC_ERROR; init returns 0
// or -1 as the reference does.
enum class SignedDistExec { CPU = 0, OpenMP = 1, GPU = 2 };

inline int signed_distance_init(const std::string& file) {
  return QUEST_signed_distance_init_serial(file.c_str());
}
// signed_distance_init(const mint::Mesh*) for a triangle mesh held as SoA
// node coordinates and connectivity, in host or device memory
inline int signed_distance_init(const double* x, const double* y, const double* z,
                                IndexType nnodes, const IndexType* connectivity, IndexType ncells) {
  return axsq_signed_distance_init_mesh(x, y, z, nnodes, connectivity, ncells);
}
inline bool signed_distance_initialized() { return QUEST_signed_distance_initialized(); }
inline void signed_distance_set_dimension(int dim) {
  QUEST_signed_distance_set_dimension(dim);
  ::axom::detail::quest_check();
}
inline void signed_distance_set_closed_surface(bool status) {
  QUEST_signed_distance_set_closed_surface(status);
  ::axom::detail::quest_check();
}
inline void signed_distance_set_compute_signs(bool computeSign) {
  QUEST_signed_distance_set_compute_signs(computeSign);
  ::axom::detail::quest_check();
}
inline void signed_distance_set_allocator(int allocatorID) {
  QUEST_signed_distance_set_allocator(allocatorID);
  ::axom::detail::quest_check();
}
inline void signed_distance_set_verbose(bool status) {
  QUEST_signed_distance_set_verbose(status);
  ::axom::detail::quest_check();
}
inline void signed_distance_use_shared_memory(bool status) {
  QUEST_signed_distance_use_shared_memory(status);
  ::axom::detail::quest_check();
}
inline void signed_distance_set_execution_space(SignedDistExec execSpace) {
  QUEST_signed_distance_set_execution_space(static_cast<int>(execSpace));
  ::axom::detail::quest_check();
}
// AXSQ_FP64 (default) or AXSQ_FP32_TRAVERSAL (no reference counterpart)
inline void signed_distance_set_precision(int fp_mode) {
  axsq_signed_distance_set_precision(fp_mode);
  ::axom::detail::quest_check();
}
inline double signed_distance_evaluate(double x, double y, double z = 0.0) {
  const double d = QUEST_signed_distance_evaluate_0(x, y, z);
  ::axom::detail::quest_check();
  return d;
}
inline double signed_distance_evaluate(double x, double y, double z, double& cp_x, double& cp_y,
                                       double& cp_z, double& n_x, double& n_y, double& n_z) {
  const double d =
      QUEST_signed_distance_evaluate_1(x, y, z, &cp_x, &cp_y, &cp_z, &n_x, &n_y, &n_z);
  ::axom::detail::quest_check();
  return d;
}
// x, y, z and phi in host or device memory
inline void signed_distance_evaluate(const double* x, const double* y, const double* z,
                                     int npoints, double* phi) {
  ::axom::detail::check(axsq_signed_distance_evaluate_batch(x, y, z, npoints, phi));
}
inline void signed_distance_get_mesh_bounds(double* lo, double* hi) {
  QUEST_signed_distance_get_mesh_bounds(lo, hi);
  ::axom::detail::quest_check();
}
inline void signed_distance_finalize() { QUEST_signed_distance_finalize(); }

}  // namespace quest
}  // namespace axom
