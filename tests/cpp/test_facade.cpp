// C++ tests of the axom_mi355x.hpp facade, written like the reference's own
// tests for this path:
//   spin/tests/spin_bvh.cpp:279-404 (findBoundingBoxes, 18 cells),
//   :519-648 (findRays, 15 cells), :792-888 (findPoints, centroid -> cell)
//   quest/tests/quest_signed_distance.cpp:88-170 (sphere l1/l2/linf)
//   spin/tests/spin_bvh.cpp:409-517, :651-790, :906-1006 (the 2-D versions)
//   quest/tests/quest_meshtester.cpp:268-305 (caved-in tetrahedron, with and
//   without degenerate triangles; meshes of quest_test_utilities.hpp:353-442)
// The BVH tests run with FloatType double and float, as the reference's do
// (spin_bvh.cpp:1560-1669).
// Exit status 0 when every check passes.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <set>
#include <vector>

#include "axom_mi355x.hpp"

static int g_fail = 0;
#define EXPECT_TRUE(c)                                                    \
  do {                                                                    \
    if (!(c)) {                                                           \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);          \
      ++g_fail;                                                           \
    }                                                                     \
  } while (0)
#define EXPECT_EQ(a, b) EXPECT_TRUE((a) == (b))
#define EXPECT_NEAR(a, b, t) EXPECT_TRUE(std::fabs((a) - (b)) <= (t))

using axom::IndexType;
using BVH = axom::spin::BVH<3, axom::MI355X_EXEC>;
using Box = axom::primal::BoundingBox3;
using Pt = axom::primal::Point3;

template <typename T>
static T* to_device(const std::vector<T>& h) {
  T* d = nullptr;
  axom::detail::hip_check(hipMalloc((void**)&d, sizeof(T) * (h.empty() ? 1 : h.size())));
  if (!h.empty())
    axom::detail::hip_check(hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return d;
}
template <typename T>
static std::vector<T> to_host(const T* d, size_t n) {
  std::vector<T> h(n);
  if (n) axom::detail::hip_check(hipMemcpy(h.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost));
  return h;
}

template <int N, typename T>
using TypesOf = axom::spin::detail::BVHTypes<N, T>;

// 3 x 3 x 3 unit cells over [0,3]^3, i fastest (mint::UniformMesh order).
template <typename BoxT = Box>
static std::vector<BoxT> grid_boxes() {
  std::vector<BoxT> b;
  for (int k = 0; k < 3; ++k)
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) {
        BoxT x;
        const int c[3] = {i, j, k};
        for (int d = 0; d < 3; ++d) {
          x.min[d] = c[d];
          x.max[d] = c[d] + 1;
        }
        b.push_back(x);
      }
  return b;
}

struct Found {
  std::vector<IndexType> off, cnt, cand;
};
template <typename Fn>
static Found run_find(IndexType n, Fn&& fn) {
  axom::Array<IndexType> off(n), cnt(n), cand;
  fn(off.view(), cnt.view(), cand);
  return Found{to_host(off.data(), n), to_host(cnt.data(), n), to_host(cand.data(), cand.size())};
}
static std::set<IndexType> set_of(const Found& f, int i) {
  return std::set<IndexType>(f.cand.begin() + f.off[i], f.cand.begin() + f.off[i] + f.cnt[i]);
}

template <typename T>
static void test_bvh_grid() {
  using Box = typename TypesOf<3, T>::BoxType;
  using Pt = typename TypesOf<3, T>::PointType;
  using BVH = axom::spin::BVH<3, axom::MI355X_EXEC, T>;
  std::vector<Box> boxes = grid_boxes<Box>();
  Box* d_boxes = to_device(boxes);
  BVH bvh;
  bvh.setScaleFactor(1.0);
  EXPECT_EQ(bvh.initialize(d_boxes, (IndexType)boxes.size()), axom::spin::BVH_BUILD_OK);
  Box b = bvh.getBounds();
  for (int d = 0; d < 3; ++d) {
    EXPECT_NEAR(b.min[d], 0.0, 1e-6);
    EXPECT_NEAR(b.max[d], 3.0, 1e-6);
  }
  // findBoundingBoxes
  std::vector<Box> q = {Box{{-1, -1, -1}, {T(2.5), T(2.5), T(1.5)}},
                        Box{{-1, -1, -1}, {T(-0.5), T(-0.5), T(-0.5)}}};
  Box* d_q = to_device(q);
  Found fb = run_find(2, [&](auto o, auto c, auto& cand) { bvh.findBoundingBoxes(o, c, cand, 2, d_q); });
  EXPECT_EQ(fb.cnt[0], 18);
  EXPECT_EQ(fb.cnt[1], 0);
  std::set<IndexType> s = set_of(fb, 0);
  for (IndexType i = 0; i < 18; ++i) EXPECT_TRUE(s.count(i) == 1);
  // findRays
  const T o[3] = {-1, -1, -1}, d1[3] = {1, 1, 1}, d2[3] = {-1, -1, -1};
  std::vector<typename TypesOf<3, T>::RayType> rays = {axom::primal::make_ray(o, d1),
                                                       axom::primal::make_ray(o, d2)};
  auto* d_rays = to_device(rays);
  Found fr = run_find(2, [&](auto oo, auto c, auto& cand) { bvh.findRays(oo, c, cand, 2, d_rays); });
  EXPECT_EQ(fr.cnt[0], 15);
  EXPECT_EQ(fr.cnt[1], 0);
  EXPECT_TRUE(set_of(fr, 0) == std::set<IndexType>({0, 1, 3, 4, 9, 10, 12, 13, 14, 16, 17, 22, 23, 25, 26}));
  // findPoints: centroids, then shifted by +10
  std::vector<Pt> c, cs;
  for (const Box& bx : boxes) {
    Pt p{{T(0.5 * (bx.min[0] + bx.max[0])), T(0.5 * (bx.min[1] + bx.max[1])),
          T(0.5 * (bx.min[2] + bx.max[2]))}};
    c.push_back(p);
    cs.push_back(Pt{{p.x[0] + 10, p.x[1] + 10, p.x[2] + 10}});
  }
  Pt* d_c = to_device(c);
  Pt* d_cs = to_device(cs);
  const IndexType n = (IndexType)c.size();
  Found fp = run_find(n, [&](auto oo, auto cc, auto& cand) { bvh.findPoints(oo, cc, cand, n, d_c); });
  for (IndexType i = 0; i < n; ++i) {
    EXPECT_EQ(fp.cnt[i], 1);
    EXPECT_EQ(fp.cand[fp.off[i]], i);
  }
  Found fs = run_find(n, [&](auto oo, auto cc, auto& cand) { bvh.findPoints(oo, cc, cand, n, d_cs); });
  for (IndexType i = 0; i < n; ++i) EXPECT_EQ(fs.cnt[i], 0);
  // initialize again on the same object: the tree is rebuilt in place over
  // boxes shifted by +10, which the shifted centroids now find
  std::vector<Box> shifted = boxes;
  for (Box& bx : shifted)
    for (int d = 0; d < 3; ++d) {
      bx.min[d] += 10;
      bx.max[d] += 10;
    }
  Box* d_shifted = to_device(shifted);
  EXPECT_EQ(bvh.initialize(d_shifted, n), axom::spin::BVH_BUILD_OK);
  Found fa = run_find(n, [&](auto oo, auto cc, auto& cand) { bvh.findPoints(oo, cc, cand, n, d_cs); });
  for (IndexType i = 0; i < n; ++i) {
    EXPECT_EQ(fa.cnt[i], 1);
    EXPECT_EQ(fa.cand[fa.off[i]], i);
  }
  // offsets of the wrong length are rejected (LinearBVH.hpp:284-285)
  bool threw = false;
  try {
    axom::Array<IndexType> off(3), cnt(n), cand;
    bvh.findPoints(off.view(), cnt.view(), cand, n, d_c);
  } catch (const axom::mi355x_error&) {
    threw = true;
  }
  EXPECT_TRUE(threw);
  for (void* p : {(void*)d_boxes, (void*)d_q, (void*)d_rays, (void*)d_c, (void*)d_cs, (void*)d_shifted})
    (void)hipFree(p);
}

// quest/tests/quest_test_utilities.hpp:51-151 (sphere), :57-87 (16^3 grid)
static void test_sdf_sphere() {
  const double R = 0.5;
  const int TR = 25, PR = 25;
  const double deg = M_PI / 180.;
  std::vector<double> x{0, 0}, y{0, 0}, z{R, -R};
  const double dphi = 180 * deg / double(PR - 1), dth = 360 * deg / double(TR - 1);
  for (int i = 0; i < TR; ++i)
    for (int j = 0; j < PR - 2; ++j) {
      const double th = i * dth, ph = j * dphi, r = R * std::sin(ph);
      x.push_back(r * std::cos(th));
      y.push_back(r * std::sin(th));
      z.push_back(R * std::cos(ph));
    }
  const int pr = PR - 2, stride = pr * TR;
  std::vector<IndexType> conn;
  for (int i = 0; i < TR; ++i) conn.insert(conn.end(), {0, (pr * (i + 1) % stride) + 2, pr * i + 2});
  for (int i = 0; i < TR; ++i) conn.insert(conn.end(), {1, (pr * (i + 1) % stride) + PR - 1, pr * i + PR - 1});
  for (int i = 0; i < TR; ++i)
    for (int j = 0; j < PR - 3; ++j) {
      const int c0 = pr * i + j + 2, c2 = ((pr * (i + 1) + j) % stride) + 3;
      conn.insert(conn.end(), {c0, c0 + 1, c2});
      conn.insert(conn.end(), {c0, c2, c2 - 1});
    }
  double *dx = to_device(x), *dy = to_device(y), *dz = to_device(z);
  IndexType* dc = to_device(conn);
  axom::quest::SignedDistance<3, axom::MI355X_EXEC> sd(dx, dy, dz, (IndexType)x.size(), dc,
                                                       (IndexType)conn.size() / 3, true, true);
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (size_t v = 0; v < x.size(); ++v) {
    const double p[3] = {x[v], y[v], z[v]};
    for (int d = 0; d < 3; ++d) {
      if (p[d] < lo[d]) lo[d] = p[d];
      if (p[d] > hi[d]) hi[d] = p[d];
    }
  }
  const int N = 16;
  std::vector<Pt> q;
  for (int d = 0; d < 3; ++d) {
    lo[d] -= 2.0;
    hi[d] += 2.0;
  }
  for (int k = 0; k < N; ++k)
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < N; ++i)
        q.push_back(Pt{{lo[0] + i * (hi[0] - lo[0]) / (N - 1), lo[1] + j * (hi[1] - lo[1]) / (N - 1),
                        lo[2] + k * (hi[2] - lo[2]) / (N - 1)}});
  Pt* dq = to_device(q);
  double* dphi_out = nullptr;
  axom::detail::hip_check(hipMalloc((void**)&dphi_out, sizeof(double) * q.size()));
  sd.computeDistances((int)q.size(), dq, dphi_out);
  std::vector<double> phi = to_host(dphi_out, q.size());
  double l1 = 0, l2 = 0, linf = 0;
  for (size_t i = 0; i < q.size(); ++i) {
    const double e = std::sqrt(q[i].x[0] * q[i].x[0] + q[i].x[1] * q[i].x[1] + q[i].x[2] * q[i].x[2]) - R;
    const double diff = phi[i] - e;
    EXPECT_NEAR(phi[i], e, 1e-2);
    l1 += std::fabs(diff);
    l2 += diff;
    linf = std::fabs(diff) > linf ? std::fabs(diff) : linf;
  }
  l2 = std::sqrt(l2);
  EXPECT_NEAR(l1, 6.7051997372579715, 1e-3);
  EXPECT_NEAR(l2, 2.5894400431865519, 1e-3);
  EXPECT_NEAR(linf, 0.00532092, 1e-3);
  EXPECT_NEAR(sd.computeDistance(q[0]), phi[0], 0.0);
  for (void* p : {(void*)dx, (void*)dy, (void*)dz, (void*)dc, (void*)dq, (void*)dphi_out}) (void)hipFree(p);
}

// The caved-in tetrahedron: triangle 0 pokes through triangles 1 and 2. With
// the two degenerate triangles appended, they are reported as degenerate; the
// BVH tester still pairs them (quest/detail/MeshTester_detail.hpp:235-289 does
// not skip them), so only the pairs among the first four are compared.
static void test_mesh_tester_caved_tet() {
  const std::vector<double> x = {2.00003, -18.213671, 4.880339, -0.000003, 13.333334};
  const std::vector<double> y = {1.00003, 4.880339, -18.213671, -0.000003, 13.333334};
  const std::vector<double> z = {18.999999, -6.666668, -6.666668, 19.999999, -6.666663};
  for (int with_degen = 0; with_degen < 2; ++with_degen) {
    std::vector<IndexType> conn = {0, 1, 2, 3, 4, 1, 3, 2, 4, 1, 4, 2};
    if (with_degen) conn.insert(conn.end(), {0, 0, 0, 3, 4, 3});
    double *dx = to_device(x), *dy = to_device(y), *dz = to_device(z);
    IndexType* dc = to_device(conn);
    std::vector<std::pair<int, int>> isect;
    std::vector<int> degen;
    axom::quest::findTriMeshIntersectionsBVH<axom::MI355X_EXEC, double>(
        dx, dy, dz, 5, dc, (IndexType)(conn.size() / 3), isect, degen);
    std::set<std::pair<int, int>> got;
    for (auto& p : isect)
      if (p.first < 4 && p.second < 4) got.insert(p);
    EXPECT_TRUE((got == std::set<std::pair<int, int>>{{0, 1}, {0, 2}}));
    EXPECT_TRUE((degen == (with_degen ? std::vector<int>{4, 5} : std::vector<int>{})));
    for (void* p : {(void*)dx, (void*)dy, (void*)dz, (void*)dc}) (void)hipFree(p);
  }
}

// spin::BVH<2>: 3 x 3 unit cells over [0,3]^2 (i fastest). The box
// [-1,2.5] x [-1,1.5] meets cells 0-5, the ray from (-1,-1) along (1,1) meets
// all but cells 2 and 6, and each cell centroid finds its own cell.
template <typename T>
static void test_bvh2d_grid() {
  using BVH2 = axom::spin::BVH<2, axom::MI355X_EXEC, T>;
  using Box2 = typename TypesOf<2, T>::BoxType;
  std::vector<Box2> cells;
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) cells.push_back(Box2{{T(i), T(j)}, {T(i + 1), T(j + 1)}});
  Box2* dcells = to_device(cells);
  BVH2 bvh;
  bvh.setScaleFactor(1.0);
  bvh.initialize(dcells, 9);
  const Box2 bb = bvh.getBounds();
  EXPECT_TRUE(bb.min[0] == 0.0 && bb.min[1] == 0.0 && bb.max[0] == 3.0 && bb.max[1] == 3.0);
  std::vector<Box2> q = {Box2{{-1, -1}, {T(2.5), T(1.5)}}, Box2{{-1, -1}, {T(-0.1), T(-0.1)}}};
  Box2* dq = to_device(q);
  {
    axom::Array<IndexType> off(2), cnt(2), cand;
    bvh.findBoundingBoxes(off.view(), cnt.view(), cand, 2, dq);
    auto c = to_host(cnt.data(), 2);
    auto ids = to_host(cand.data(), (size_t)cand.size());
    EXPECT_TRUE(c[0] == 6 && c[1] == 0);
    EXPECT_TRUE((std::set<IndexType>(ids.begin(), ids.end()) == std::set<IndexType>{0, 1, 2, 3, 4, 5}));
  }
  {
    const T o[2] = {-1, -1}, d0[2] = {1, 1}, d1[2] = {-1, -1};
    std::vector<typename TypesOf<2, T>::RayType> rays = {axom::primal::make_ray2(o, d0),
                                                         axom::primal::make_ray2(o, d1)};
    auto* dr = to_device(rays);
    axom::Array<IndexType> off(2), cnt(2), cand;
    bvh.findRays(off.view(), cnt.view(), cand, 2, dr);
    auto c = to_host(cnt.data(), 2);
    auto ids = to_host(cand.data(), (size_t)cand.size());
    EXPECT_TRUE(c[0] == 7 && c[1] == 0);
    EXPECT_TRUE((std::set<IndexType>(ids.begin(), ids.end()) == std::set<IndexType>{0, 1, 3, 4, 5, 7, 8}));
    (void)hipFree(dr);
  }
  {
    std::vector<typename TypesOf<2, T>::PointType> pts;
    for (auto& c : cells)
      pts.push_back({{T(0.5 * (c.min[0] + c.max[0])), T(0.5 * (c.min[1] + c.max[1]))}});
    auto* dp = to_device(pts);
    axom::Array<IndexType> off(9), cnt(9), cand;
    bvh.findPoints(off.view(), cnt.view(), cand, 9, dp);
    auto c = to_host(cnt.data(), 9);
    auto ids = to_host(cand.data(), (size_t)cand.size());
    for (int k = 0; k < 9; ++k) EXPECT_EQ(c[k], 1);
    for (int k = 0; k < 9 && k < (int)ids.size(); ++k) EXPECT_EQ(ids[k], k);
    (void)hipFree(dp);
  }
  (void)hipFree(dcells);
  (void)hipFree(dq);
}

// BVH::getTraverser in a caller's kernel (spin/policy/LinearBVH.hpp:57-109):
// a box query with the overlap predicate must reach the same leaves in the
// same order as findBoundingBoxes; a point query (nearer child first) the same
// set as findPoints.
constexpr int kTravMax = 64;
template <typename Trav, typename Box = typename Trav::BoxType>
__global__ void k_trav_boxes(Trav tr, const Box* q, int n, int* counts, int* cand,
                             int* complete = nullptr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = 0;
  const bool done = tr.traverse_tree(
      q[i],
      [&](std::int32_t pos, const std::int32_t* ids) {
        if (c < kTravMax) cand[i * kTravMax + c] = ids[pos];
        ++c;
      },
      [](const Box& a, const Box& b) {
        for (int d = 0; d < 3; ++d)
          if (a.max[d] < b.min[d] || a.min[d] > b.max[d]) return false;
        return true;
      });
  counts[i] = c;
  if (complete) complete[i] = done ? 1 : 0;
}
template <typename Trav, typename Box = typename Trav::BoxType,
          typename Pt = typename Trav::PointType>
__global__ void k_trav_points(Trav tr, const Pt* q, int n, int* counts, int* cand) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = 0;
  tr.traverse_tree(
      q[i],
      [&](std::int32_t pos, const std::int32_t* ids) {
        if (c < kTravMax) cand[i * kTravMax + c] = ids[pos];
        ++c;
      },
      [](const Pt& p, const Box& b) {
        for (int d = 0; d < 3; ++d)
          if (p.x[d] < b.min[d] || p.x[d] > b.max[d]) return false;
        return true;
      });
  counts[i] = c;
}

template <typename T>
static void test_traverser_matches_find() {
  using Box = typename TypesOf<3, T>::BoxType;
  using Pt = typename TypesOf<3, T>::PointType;
  using BVH = axom::spin::BVH<3, axom::MI355X_EXEC, T>;
  // 4,000 boxes and 3,000 box / point queries from a fixed LCG
  unsigned long long st = 12345;
  auto rnd = [&]() {
    st = st * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(st >> 11) / 9007199254740992.0;
  };
  std::vector<Box> boxes(4000);
  for (Box& b : boxes)
    for (int d = 0; d < 3; ++d) {
      const double c = rnd(), h = 0.01 + 0.04 * rnd();
      b.min[d] = T(c - h);
      b.max[d] = T(c + h);
    }
  Box* d_boxes = to_device(boxes);
  BVH bvh;
  bvh.initialize(d_boxes, (IndexType)boxes.size());
  const int n = 3000;
  std::vector<Box> qb(n);
  std::vector<Pt> qp(n);
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      const double c = rnd(), h = 0.02 * rnd();
      qb[i].min[d] = T(c - h);
      qb[i].max[d] = T(c + h);
      qp[i].x[d] = T(rnd());
    }
  Box* d_qb = to_device(qb);
  Pt* d_qp = to_device(qp);
  int *d_cnt = nullptr, *d_cand = nullptr;
  axom::detail::hip_check(hipMalloc((void**)&d_cnt, sizeof(int) * n));
  axom::detail::hip_check(hipMalloc((void**)&d_cand, sizeof(int) * n * kTravMax));
  const auto tr = bvh.getTraverser();
  // boxes: same leaves, same order
  k_trav_boxes<<<(n + 255) / 256, 256>>>(tr, d_qb, n, d_cnt, d_cand);
  axom::detail::hip_check(hipGetLastError());
  axom::detail::hip_check(hipDeviceSynchronize());
  std::vector<int> cnt = to_host(d_cnt, n), cand = to_host(d_cand, (size_t)n * kTravMax);
  Found fb = run_find(n, [&](auto o, auto c, auto& cd) { bvh.findBoundingBoxes(o, c, cd, n, d_qb); });
  long total = 0;
  for (int i = 0; i < n; ++i) {
    EXPECT_EQ(cnt[i], fb.cnt[i]);
    if (cnt[i] != fb.cnt[i] || cnt[i] > kTravMax) continue;
    total += cnt[i];
    for (int k = 0; k < cnt[i]; ++k) EXPECT_EQ(cand[i * kTravMax + k], fb.cand[fb.off[i] + k]);
  }
  EXPECT_TRUE(total > 1000);
  // points: the same set (the nearer child is visited first)
  k_trav_points<<<(n + 255) / 256, 256>>>(tr, d_qp, n, d_cnt, d_cand);
  axom::detail::hip_check(hipGetLastError());
  axom::detail::hip_check(hipDeviceSynchronize());
  cnt = to_host(d_cnt, n);
  cand = to_host(d_cand, (size_t)n * kTravMax);
  Found fp = run_find(n, [&](auto o, auto c, auto& cd) { bvh.findPoints(o, c, cd, n, d_qp); });
  for (int i = 0; i < n; ++i) {
    EXPECT_EQ(cnt[i], fp.cnt[i]);
    if (cnt[i] != fp.cnt[i] || cnt[i] > kTravMax) continue;
    std::set<IndexType> a(cand.begin() + i * kTravMax, cand.begin() + i * kTravMax + cnt[i]);
    EXPECT_TRUE(a == set_of(fp, i));
  }
  for (void* p : {(void*)d_boxes, (void*)d_qb, (void*)d_qp, (void*)d_cnt, (void*)d_cand})
    (void)hipFree(p);
}

// A tree whose walks push 37 entries: Morton codes that differ in successively
// lower bits (a chain of nodes whose far child is a small subtree) and 256
// copies of code 0 (split by index into a balanced subtree), each box
// stretched to a common corner so that a box query around everything passes
// every node. A 16-entry traverser stops early and says so, having made a
// prefix of the leaf actions; the 64-entry one finishes with
// findBoundingBoxes's order.
static void test_traverser_reports_a_full_stack() {
  auto cell = [](int qx, int qy, int qz) {
    return std::vector<double>{(qx + 0.5) / 1024, (qy + 0.5) / 1024, (qz + 0.5) / 1024};
  };
  std::vector<std::vector<double>> cents(256, cell(0, 0, 0));
  for (int k = 0; k < 30; ++k) {
    int q[3] = {0, 0, 0};
    q[k % 3] = 1 << (k / 3);
    cents.push_back(cell(q[0], q[1], q[2]));
    cents.push_back(cell(q[0], q[1], q[2]));
  }
  cents.push_back(cell(1023, 1023, 1023));
  const double corner = 0.25 / 1024, h = 1e-6;
  std::vector<Box> boxes;
  for (const auto& c : cents) {
    Box b;
    for (int d = 0; d < 3; ++d) {
      b.min[d] = std::min(c[d], corner) - h;
      b.max[d] = std::max(c[d], corner) + h;
    }
    boxes.push_back(b);
  }
  Box* d_boxes = to_device(boxes);
  BVH bvh;
  bvh.initialize(d_boxes, (IndexType)boxes.size());
  const int n = 64;
  std::vector<Box> qb(n, Box{{-1.0, -1.0, -1.0}, {2.0, 2.0, 2.0}});
  Box* d_qb = to_device(qb);
  const int cap = 1024;
  int *d_cnt = nullptr, *d_cand = nullptr, *d_done = nullptr;
  axom::detail::hip_check(hipMalloc((void**)&d_cnt, sizeof(int) * n));
  axom::detail::hip_check(hipMalloc((void**)&d_cand, sizeof(int) * n * kTravMax));
  axom::detail::hip_check(hipMalloc((void**)&d_done, sizeof(int) * n));
  Found fb = run_find(n, [&](auto o, auto c, auto& cd) { bvh.findBoundingBoxes(o, c, cd, n, d_qb); });
  EXPECT_EQ(fb.cnt[0], (IndexType)boxes.size());
  (void)cap;
  // 16 entries: stops early, reports it; the leaves it reached begin the full order
  k_trav_boxes<<<1, n>>>(bvh.getTraverserWithStack<16>(), d_qb, n, d_cnt, d_cand, d_done);
  axom::detail::hip_check(hipGetLastError());
  axom::detail::hip_check(hipDeviceSynchronize());
  std::vector<int> cnt = to_host(d_cnt, n), done = to_host(d_done, n),
                   cand = to_host(d_cand, (size_t)n * kTravMax);
  for (int i = 0; i < n; ++i) {
    EXPECT_EQ(done[i], 0);
    EXPECT_TRUE(cnt[i] < fb.cnt[i]);
    for (int k = 0; k < std::min(cnt[i], kTravMax); ++k)
      EXPECT_EQ(cand[i * kTravMax + k], fb.cand[fb.off[i] + k]);
  }
  // 64 entries: the whole walk
  k_trav_boxes<<<1, n>>>(bvh.getTraverser(), d_qb, n, d_cnt, d_cand, d_done);
  axom::detail::hip_check(hipGetLastError());
  axom::detail::hip_check(hipDeviceSynchronize());
  cnt = to_host(d_cnt, n);
  done = to_host(d_done, n);
  cand = to_host(d_cand, (size_t)n * kTravMax);
  for (int i = 0; i < n; ++i) {
    EXPECT_EQ(done[i], 1);
    EXPECT_EQ(cnt[i], fb.cnt[i]);
    for (int k = 0; k < std::min(cnt[i], kTravMax); ++k)
      EXPECT_EQ(cand[i * kTravMax + k], fb.cand[fb.off[i] + k]);
  }
  for (void* p : {(void*)d_boxes, (void*)d_qb, (void*)d_cnt, (void*)d_cand, (void*)d_done})
    (void)hipFree(p);
}

// BVH::setAllocatorID: candidate arrays in managed memory, read on the host.
static void test_allocator_and_vtk() {
  std::vector<Box> boxes = grid_boxes();
  Box* d_boxes = to_device(boxes);
  BVH bvh;
  bvh.setAllocatorID(AXSQ_ALLOC_MANAGED);
  EXPECT_EQ(bvh.getAllocatorID(), AXSQ_ALLOC_MANAGED);
  bvh.initialize(d_boxes, (IndexType)boxes.size());
  std::vector<Pt> pts;
  for (const Box& b : boxes)
    pts.push_back(Pt{{0.5 * (b.min[0] + b.max[0]), 0.5 * (b.min[1] + b.max[1]),
                      0.5 * (b.min[2] + b.max[2])}});
  Pt* d_pts = to_device(pts);
  const IndexType n = (IndexType)pts.size();
  axom::Array<IndexType> off(n), cnt(n), cand;
  bvh.findPoints(off.view(), cnt.view(), cand, n, d_pts);
  axom::detail::hip_check(hipDeviceSynchronize());
  EXPECT_EQ(cand.size(), n);
  for (IndexType i = 0; i < n && cand.size() == n; ++i) EXPECT_EQ(cand.data()[i], i);  // host read
  bool threw = false;
  try {
    bvh.setAllocatorID(42);
  } catch (const axom::mi355x_error&) {
    threw = true;
  }
  EXPECT_TRUE(threw);
  // BVH::writeVtkFile: 1 root + 2 per inner node boxes, 8 points each
  bvh.writeVtkFile("test_facade_bvh.vtk");
  FILE* f = std::fopen("test_facade_bvh.vtk", "r");
  EXPECT_TRUE(f != nullptr);
  if (f) {
    char line[128] = {0};
    EXPECT_TRUE(std::fgets(line, sizeof(line), f) != nullptr);
    EXPECT_EQ(std::string(line), std::string("# vtk DataFile Version 3.0\n"));
    int npts = 0;
    while (std::fgets(line, sizeof(line), f))
      if (std::sscanf(line, "POINTS %d double", &npts) == 1) break;
    EXPECT_EQ(npts, 8 * (1 + 2 * (n - 1)));
    std::fclose(f);
    std::remove("test_facade_bvh.vtk");
  }
  for (void* p : {(void*)d_boxes, (void*)d_pts}) (void)hipFree(p);
}

// quest::SignedDistance single points and the signed_distance_* interface
// (signed_distance.hpp:88-319) over an octahedron |x|+|y|+|z| = 1.
static void test_signed_distance_interface() {
  std::vector<double> x{1, -1, 0, 0, 0, 0}, y{0, 0, 1, -1, 0, 0}, z{0, 0, 0, 0, 1, -1};
  std::vector<IndexType> conn{0, 2, 4, 2, 1, 4, 1, 3, 4, 3, 0, 4,
                              2, 0, 5, 1, 2, 5, 3, 1, 5, 0, 3, 5};
  double *dx = to_device(x), *dy = to_device(y), *dz = to_device(z);
  IndexType* dc = to_device(conn);
  axom::quest::SignedDistance<3, axom::MI355X_EXEC> sd(dx, dy, dz, 6, dc, 8, true, true);
  // computeDistance(pt, cp, normal) against the batched form
  std::vector<Pt> q{Pt{{2.0, 2.0, 2.0}}, Pt{{0.1, 0.0, 0.0}}, Pt{{0.0, 0.0, 3.0}}};
  Pt* dq = to_device(q);
  double* dd = nullptr;
  Pt* dcp = nullptr;
  axom::primal::Vector3* dn = nullptr;
  axom::detail::hip_check(hipMalloc((void**)&dd, sizeof(double) * 3));
  axom::detail::hip_check(hipMalloc((void**)&dcp, sizeof(Pt) * 3));
  axom::detail::hip_check(hipMalloc((void**)&dn, sizeof(axom::primal::Vector3) * 3));
  sd.computeDistances(3, dq, dd, dcp, dn);
  std::vector<double> d = to_host(dd, 3);
  std::vector<Pt> cp = to_host(dcp, 3);
  std::vector<axom::primal::Vector3> nv = to_host(dn, 3);
  for (int i = 0; i < 3; ++i) {
    Pt c;
    axom::primal::Vector3 nn;
    EXPECT_EQ(sd.computeDistance(q[i], c, nn), d[i]);
    EXPECT_EQ(sd.computeDistance(q[i]), d[i]);
    for (int k = 0; k < 3; ++k) {
      EXPECT_EQ(c.x[k], cp[i].x[k]);
      EXPECT_EQ(nn.x[k], nv[i].x[k]);
    }
  }
  EXPECT_NEAR(d[0], (6.0 - 1.0) / std::sqrt(3.0), 1e-15);  // outside, face (1,1,1)
  EXPECT_NEAR(d[2], 2.0, 1e-15);                           // above the vertex (0,0,1)
  EXPECT_TRUE(d[1] < 0.0);                                 // inside
  // the process-wide query, initialized from host arrays
  namespace q_ = axom::quest;
  q_::signed_distance_finalize();
  q_::signed_distance_set_closed_surface(true);
  q_::signed_distance_set_execution_space(q_::SignedDistExec::GPU);
  EXPECT_EQ(q_::signed_distance_init(x.data(), y.data(), z.data(), 6, conn.data(), 8), 0);
  EXPECT_TRUE(q_::signed_distance_initialized());
  bool threw = false;
  try {
    q_::signed_distance_set_compute_signs(false);
  } catch (const axom::mi355x_error&) {
    threw = true;
  }
  EXPECT_TRUE(threw);
  double cx, cy, cz, nx, ny, nz;
  EXPECT_EQ(q_::signed_distance_evaluate(2.0, 2.0, 2.0), d[0]);
  EXPECT_EQ(q_::signed_distance_evaluate(0.1, 0.0, 0.0, cx, cy, cz, nx, ny, nz), d[1]);
  EXPECT_TRUE(cx == cp[1].x[0] && cy == cp[1].x[1] && cz == cp[1].x[2]);
  EXPECT_TRUE(nx == nv[1].x[0] && ny == nv[1].x[1] && nz == nv[1].x[2]);
  std::vector<double> hx{2.0, 0.1, 0.0}, hy{2.0, 0.0, 0.0}, hz{2.0, 0.0, 3.0}, phi(3, 0.0);
  q_::signed_distance_evaluate(hx.data(), hy.data(), hz.data(), 3, phi.data());  // host arrays
  for (int i = 0; i < 3; ++i) EXPECT_EQ(phi[i], d[i]);
  double lo[3], hi[3];
  q_::signed_distance_get_mesh_bounds(lo, hi);
  for (int k = 0; k < 3; ++k) {
    EXPECT_EQ(lo[k], -1.0);
    EXPECT_EQ(hi[k], 1.0);
  }
  q_::signed_distance_finalize();
  EXPECT_TRUE(!q_::signed_distance_initialized());
  threw = false;
  try {
    (void)q_::signed_distance_evaluate(0.0, 0.0, 0.0);
  } catch (const axom::mi355x_error&) {
    threw = true;
  }
  EXPECT_TRUE(threw);
  // STLReader (quest_stl_reader.cpp:99-140): the reference's one-triangle file
  {
    FILE* f = std::fopen("test_facade_triangle.stl", "w");
    std::fputs("solid triangle\n\t facet normal 0.0 0.0 1.0\n\t\t outer loop\n"
               "\t\t\t vertex 0.0 0.0 0.0\n\t\t\t vertex 1.0 0.0 0.0\n"
               "\t\t\t vertex 0.0 1.0 0.0\n\t\t endloop\n\t endfacet\n"
               "endsolid triangle\n", f);
    std::fclose(f);
    axom::quest::STLReader reader;
    reader.setFileName("test_facade_triangle.stl");
    EXPECT_EQ(reader.read(), 0);
    EXPECT_EQ(reader.getNumNodes(), 3);
    EXPECT_EQ(reader.getNumFaces(), 1);
    std::vector<double> mx, my, mz;
    std::vector<IndexType> mc;
    reader.getMesh(mx, my, mz, mc);
    EXPECT_TRUE(mx == std::vector<double>({0.0, 1.0, 0.0}));
    EXPECT_TRUE(my == std::vector<double>({0.0, 0.0, 1.0}));
    EXPECT_TRUE(mz == std::vector<double>({0.0, 0.0, 0.0}));
    EXPECT_TRUE(mc == std::vector<IndexType>({0, 1, 2}));
    std::remove("test_facade_triangle.stl");
    axom::quest::STLReader missing("foo.stl");
    EXPECT_TRUE(missing.read() != 0);
  }
  for (void* p : {(void*)dx, (void*)dy, (void*)dz, (void*)dc, (void*)dq, (void*)dd, (void*)dcp,
                  (void*)dn})
    (void)hipFree(p);
}

int main() {
  test_bvh_grid<double>();
  test_bvh_grid<float>();
  test_traverser_reports_a_full_stack();
  test_allocator_and_vtk();
  test_signed_distance_interface();
  test_traverser_matches_find<double>();
  test_traverser_matches_find<float>();
  test_bvh2d_grid<double>();
  test_bvh2d_grid<float>();
  test_sdf_sphere();
  test_mesh_tester_caved_tet();
  std::printf("%s (%d failures)\n", g_fail ? "FAILED" : "PASSED", g_fail);
  return g_fail ? 1 : 0;
}
