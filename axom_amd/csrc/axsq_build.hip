// axsq_build.hip -- linear BVH construction on gfx950.
//
// Pipeline (reference: src/axom/spin/internal/linear_bvh/build_radix_tree.hpp:579-611
// and src/axom/spin/policy/LinearBVH.hpp:193-269):
//   k_transform   scale every box about its centroid (+ padding of 0/1 boxes)
//                 and reduce block-wise the bounds of the valid boxes
//   k_bounds      finish the bounds; Morton normalisation constants
//   k_mcodes      30-bit Morton code of each scaled-box centroid
//   radix sort    stable LSD sort of (code, id)
//   k_reorder     scaled leaf boxes in leaf (sorted) order
//   k_karras      Karras 2012 hierarchy over the sorted codes
//   k_refit_*     bottom-up boxes; the second thread to reach a node goes on
//                 (hand-offs kept inside workgroups by tiers of leaf spans)
//   k_emit        128-byte FP64 nodes (+ 64-byte FP32 nodes, rounded outward)
// Up to kSmallSortMax boxes the same tree takes eight launches: k_transform;
// k_mcodes_small (bounds + codes + the sort's first histogram); the sort's
// passes, the last writing the leaves in tree order; k_seg_local (the
// hierarchy by wide searches beside a balanced segment tree over each 512
// leaves in LDS, whose unions give the child boxes of every node inside
// those leaves); k_emit_seg (the few nodes that reach across) -- no climb.

#include <mutex>

#include "axsq_internal.hpp"

namespace axsq {

namespace {

struct BuildParams {
  double bounds[6];
  double minc[3];
  double inv[3];
  int32_t quirky;  // some leaf box has a coordinate with min -0.0 and max +0.0 (or the reverse)
  int32_t pad;
};

__device__ __forceinline__ void load_box(const double* __restrict__ p, double b[6]) {
  const double2* q = reinterpret_cast<const double2*>(p);
  if ((reinterpret_cast<uintptr_t>(p) & 15) == 0) {
    const double2 a = q[0], c = q[1], d = q[2];
    b[0] = a.x; b[1] = a.y; b[2] = c.x; b[3] = c.y; b[4] = d.x; b[5] = d.y;
  } else {
    for (int k = 0; k < 6; ++k) b[k] = p[k];
  }
}

__device__ __forceinline__ void store_box(double* __restrict__ p, const double b[6]) {
  double2* q = reinterpret_cast<double2*>(p);
  q[0] = make_double2(b[0], b[1]);
  q[1] = make_double2(b[2], b[3]);
  q[2] = make_double2(b[4], b[5]);
}

__device__ __forceinline__ void invalid_box(double b[6], double big = kDblMax) {
  b[0] = b[1] = b[2] = big;
  b[3] = b[4] = b[5] = -big;
}

// Per-block partial of k_transform: min/max of the block's valid scaled boxes
// (6 doubles), then 1.0 if one of its boxes has a coordinate whose min and
// max are equal zeros of opposite signs (else 0.0), then padding.
constexpr int kPartStride = 8;

// A coordinate whose min and max compare equal but differ in their bits
// (-0.0 and +0.0). Box unions (BoundingBox::addBox) over leaves that hold one
// can depend on the grouping, not only on the order, of the leaves.
__device__ __forceinline__ bool box_signed_zero_pair(const double b[6]) {
  bool q = false;
  for (int d = 0; d < 3; ++d)
    q |= (b[d] == b[3 + d]) && (__double_as_longlong(b[d]) != __double_as_longlong(b[3 + d]));
  return q;
}

// Per block: scaled boxes out + min/max of the valid ones. A block with no
// valid box reports an invalid partial.
// The same launch clears zero[0, nzero) (the sort's histograms).
// F32: a float tree; `in` holds float boxes of 2 * in_dims floats (min, max),
// widened to doubles (z = [0, 0] in 2-D), the padding box is
// BoundingBox<float>'s invalid box and the scaling is in float.
template <bool F32>
__global__ __launch_bounds__(kBlock) void k_transform(const void* __restrict__ in_v, int in_dims,
                                                       int64_t n_in, int64_t n, double scale,
                                                       double* __restrict__ out,
                                                       double* __restrict__ partials,
                                                       uint32_t* __restrict__ zero, int64_t nzero) {
  __shared__ double s_red[6][kBlock];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t j = i; j < nzero; j += (int64_t)gridDim.x * blockDim.x) zero[j] = 0u;
  double acc[6];
  invalid_box(acc);
  int quirky = 0;
  if (i < n) {
    double b[6];
    if (n_in > 1 || (n_in == 1 && i == 0)) {
      if constexpr (F32) {
        const float* p = static_cast<const float*>(in_v) + 2 * in_dims * i;
        for (int d = 0; d < 3; ++d) {
          b[d] = d < in_dims ? (double)p[d] : 0.0;
          b[3 + d] = d < in_dims ? (double)p[in_dims + d] : 0.0;
        }
      } else {
        load_box(static_cast<const double*>(in_v) + 6 * i, b);
      }
    } else {
      invalid_box(b, F32 ? (double)kFltMax : kDblMax);
    }
    if constexpr (F32)
      box_scale_f32(b, (float)scale);
    else
      box_scale(b, scale);
    store_box(out + 6 * i, b);
    if (box_valid(b))
      for (int k = 0; k < 6; ++k) acc[k] = b[k];
    quirky = box_signed_zero_pair(b);
  }
  quirky = __syncthreads_or(quirky);
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
  if (threadIdx.x == 0) {
    double* p = partials + kPartStride * (int64_t)blockIdx.x;
    for (int k = 0; k < 6; ++k) p[k] = s_red[k][0];
    p[6] = quirky ? 1.0 : 0.0;
    p[7] = 0.0;
  }
}

// Union of the valid scaled boxes; with none valid, the SEQ reduce
// (addBox from an invalid box, build_radix_tree.hpp:134-142) leaves the last
// box. Then inv_extent as in get_mcodes (:154-163), in float for a float
// tree (f32: extent and 1 / extent rounded to float, the zero test's
// threshold 1e-8 as a float). Run by a whole 1024-thread workgroup, which
// finds the partials' min/max in one fixed order; thread 0 returns the
// parameters (and whether any partial flags a signed-zero pair).
__device__ void finish_params(const double* __restrict__ partials, int64_t nparts,
                              const double* __restrict__ boxes, int64_t n, int f32,
                              double (*s_red)[1024], BuildParams& out, int& quirky) {
  const int t = threadIdx.x;
  double acc[6];
  invalid_box(acc);
  int q = 0;
  for (int64_t p = t; p < nparts; p += 1024) {
    const double* pp = partials + kPartStride * p;
    for (int k = 0; k < 3; ++k) {
      const double a = pp[k];
      if (a < acc[k]) acc[k] = a;
      const double b = pp[3 + k];
      if (b > acc[3 + k]) acc[3 + k] = b;
    }
    q |= (pp[6] != 0.0) ? 1 : 0;
  }
  quirky = __syncthreads_or(q);
  for (int k = 0; k < 6; ++k) s_red[k][t] = acc[k];
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (t < off) {
      for (int k = 0; k < 3; ++k) {
        const double a = s_red[k][t], b = s_red[k][t + off];
        s_red[k][t] = (b < a) ? b : a;
        const double c = s_red[3 + k][t], d = s_red[3 + k][t + off];
        s_red[3 + k][t] = (d > c) ? d : c;
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    double b[6];
    for (int k = 0; k < 6; ++k) b[k] = s_red[k][0];
    if (!box_valid(b)) load_box(boxes + 6 * (n - 1), b);
    for (int k = 0; k < 6; ++k) out.bounds[k] = b[k];
    for (int d = 0; d < 3; ++d) {
      out.minc[d] = b[d];
      if (f32) {
        const float ext = (float)b[3 + d] - (float)b[d];
        out.inv[d] = (fabsf(ext - 0.0f) <= (float)1.0e-8) ? 0.0 : (double)(1.0f / ext);
      } else {
        const double ext = b[3 + d] - b[d];
        out.inv[d] = near_eq(ext, 0.0, 1.0e-8) ? 0.0 : 1.0 / ext;
      }
    }
  }
}

// prm: device copy; host_out (pinned host memory): a copy for the host.
__global__ __launch_bounds__(1024) void k_bounds(const double* __restrict__ partials,
                                                  int64_t nparts, const double* __restrict__ boxes,
                                                  int64_t n, BuildParams* __restrict__ prm,
                                                  BuildParams* __restrict__ host_out, int f32) {
  __shared__ double s_red[6][1024];
  BuildParams p;
  int quirky;
  finish_params(partials, nparts, boxes, n, f32, s_red, p, quirky);
  if (threadIdx.x == 0) {
    p.quirky = quirky;
    p.pad = 0;
    *prm = p;
    *host_out = p;
  }
}

// Morton code of scaled box i's centroid (f32: the normalised centroid in float).
__device__ __forceinline__ uint32_t box_code(const double b[6], const double minc_in[3],
                                             const double inv_in[3], int dims, int f32) {
  double c[3], minc[3], inv[3];
  for (int d = 0; d < 3; ++d) {
    c[d] = 0.5 * (b[d] + b[3 + d]);
    minc[d] = minc_in[d];
    inv[d] = inv_in[d];
  }
  if (f32) {  // (c - min) * inv in float, then codes of that point over [0, 1]
    mcode_coords_f32(b, minc, inv, c);
    for (int d = 0; d < 3; ++d) {
      minc[d] = 0.0;
      inv[d] = 1.0;
    }
  }
  return (dims == 2) ? morton2_of(c, minc, inv) : morton_of(c, minc, inv);
}

__global__ __launch_bounds__(kBlock) void k_mcodes(const double* __restrict__ boxes, int64_t n,
                                                    const BuildParams* __restrict__ prm, int dims,
                                                    int f32, uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double b[6];
  load_box(boxes + 6 * i, b);
  keys[i] = box_code(b, prm->minc, prm->inv, dims, f32);
  vals[i] = (uint32_t)i;
}

// Small builds: the bounds and constants (every workgroup finds the same
// ones from the transform's partials; workgroup 0 writes them to host_out,
// pinned host memory), then the
// code of box i = 1024 * workgroup + thread, and the sort's first digit
// histogram: workgroup g is sort tile g, so its counts go straight to
// hist0[g * 256 + digit].
__global__ __launch_bounds__(1024) void k_mcodes_small(const double* __restrict__ partials,
                                                        int64_t nparts,
                                                        const double* __restrict__ boxes,
                                                        int64_t n, int dims, int f32,
                                                        BuildParams* __restrict__ host_out,
                                                        uint32_t* __restrict__ keys,
                                                        uint32_t* __restrict__ vals,
                                                        uint32_t* __restrict__ hist0) {
  static_assert(kSmallSortTile == 1024, "one sort tile per workgroup");
  __shared__ double s_red[6][1024];
  __shared__ double s_prm[6];
  __shared__ uint32_t s_h[256];
  const int t = threadIdx.x;
  if (t < 256) s_h[t] = 0u;
  BuildParams p;
  int quirky;
  finish_params(partials, nparts, boxes, n, f32, s_red, p, quirky);
  if (t == 0) {
    if (blockIdx.x == 0) {
      p.quirky = quirky;
      p.pad = 0;
      *host_out = p;
    }
    for (int d = 0; d < 3; ++d) {
      s_prm[d] = p.minc[d];
      s_prm[3 + d] = p.inv[d];
    }
  }
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * 1024 + t;
  if (i < n) {
    double b[6];
    load_box(boxes + 6 * i, b);
    const double minc[3] = {s_prm[0], s_prm[1], s_prm[2]};
    const double inv[3] = {s_prm[3], s_prm[4], s_prm[5]};
    const uint32_t code = box_code(b, minc, inv, dims, f32);
    keys[i] = code;
    vals[i] = (uint32_t)i;
    atomicAdd(&s_h[code & 255u], 1u);
  }
  __syncthreads();
  if (t < 256) hist0[(int64_t)blockIdx.x * 256 + t] = s_h[t];
}

__global__ __launch_bounds__(kBlock) void k_reorder(const double* __restrict__ boxes,
                                                     const uint32_t* __restrict__ perm,
                                                     const uint32_t* __restrict__ codes,
                                                     int64_t n, double* __restrict__ leaf_boxes,
                                                     int32_t* __restrict__ leaf_ids,
                                                     uint32_t* __restrict__ mcodes) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t src = perm[p];
  double b[6];
  load_box(boxes + 6 * (int64_t)src, b);
  store_box(leaf_boxes + 6 * p, b);
  leaf_ids[p] = (int32_t)src;
  mcodes[p] = codes[p];
}

// build_tree, build_radix_tree.hpp:290-384: inner node i's range, split and
// links.
__device__ __forceinline__ void karras_node(const uint32_t* __restrict__ mc, int32_t inner,
                                            int32_t i, int32_t* __restrict__ lchild,
                                            int32_t* __restrict__ rchild,
                                            int32_t* __restrict__ parents,
                                            int32_t* __restrict__ range) {
  const int32_t d =
      (karras_delta(i, i + 1, inner, mc) - karras_delta(i, i - 1, inner, mc)) < 0 ? -1 : 1;
  const int32_t dmin = karras_delta(i, i - d, inner, mc);
  int32_t lmax = 2;
  while (karras_delta(i, i + lmax * d, inner, mc) > dmin) lmax *= 2;
  int32_t l = 0;
  for (int32_t s = lmax / 2; s >= 1; s /= 2)
    if (karras_delta(i, i + (l + s) * d, inner, mc) > dmin) l += s;
  const int32_t j = i + l * d;
  const int32_t dnode = karras_delta(i, j, inner, mc);
  int32_t s = 0;
  double div = 2.0;
  int32_t step = (int32_t)ceil((double)(float)l / div);
  for (;;) {
    if (karras_delta(i, i + (s + step) * d, inner, mc) > dnode) s += step;
    if (step == 1) break;
    div *= 2;
    step = (int32_t)ceil((double)(float)l / div);
  }
  const int32_t split = i + s * d + (d < 0 ? d : 0);
  const int32_t lo = (i < j) ? i : j;
  const int32_t hi = (i < j) ? j : i;
  range[3 * i] = lo;
  range[3 * i + 1] = hi;
  range[3 * i + 2] = split;
  if (lo == split) {
    parents[split + inner] = i;
    lchild[i] = split + inner;
  } else {
    parents[split] = i;
    lchild[i] = split;
  }
  if (hi == split + 1) {
    parents[split + inner + 1] = i;
    rchild[i] = split + inner + 1;
  } else {
    parents[split + 1] = i;
    rchild[i] = split + 1;
  }
  if (i == 0) parents[0] = -1;
}

// The same node with several loads in flight per step, for trees of at most
// kSmallSortMax leaves. delta(i, i + x * d) does not increase with x, so the
// range end and the split are the last positions that pass their tests: the
// reference's doubling search and binary searches find exactly those (the
// split search's float-rounded steps are exact below 2^24), and so do
// 8-way searches that test 7 positions at a time.
__device__ __forceinline__ void karras_node_wide(const uint32_t* __restrict__ mc, int32_t inner,
                                                 int32_t i, int32_t* __restrict__ lchild,
                                                 int32_t* __restrict__ rchild,
                                                 int32_t* __restrict__ parents,
                                                 int32_t* __restrict__ range, int32_t& lo_out,
                                                 int32_t& hi_out, int32_t& split_out) {
  const uint32_t ci = mc[i];
  auto delta = [&](int32_t b) -> int32_t {  // karras_delta(i, b)
    const bool oor = (b < 0 || b > inner);
    const int32_t bb = oor ? 0 : b;
    uint32_t ex = ci ^ mc[bb];
    const bool tie = (ex == 0);
    if (tie) ex = (uint32_t)i ^ (uint32_t)bb;
    int32_t cnt = ex ? __clz((int)ex) : 32;
    if (tie) cnt += 32;
    return oor ? -1 : cnt;
  };
  // the last x in [lo, hi] with delta(i, i + x * d) > thr, given that lo passes
  auto last_pass = [&](int32_t lo, int32_t hi, int32_t d, int32_t thr) -> int32_t {
    while (hi > lo) {
      const int64_t span = hi - lo;
      int32_t p[7];
      bool v[7];
#pragma unroll
      for (int k = 0; k < 7; ++k) p[k] = lo + (int32_t)((span * (k + 1) + 7) / 8);
#pragma unroll
      for (int k = 0; k < 7; ++k) v[k] = delta(i + p[k] * d) > thr;
      int32_t nlo = lo, nhi = hi;
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        if (v[k])
          nlo = p[k] > nlo ? p[k] : nlo;
        else
          nhi = p[k] - 1 < nhi ? p[k] - 1 : nhi;
      }
      lo = nlo;
      hi = nhi;
    }
    return lo;
  };
  const int32_t dp = delta(i + 1), dm = delta(i - 1);
  const int32_t d = (dp - dm) < 0 ? -1 : 1;
  const int32_t dmin = d > 0 ? dm : dp;
  // doubling search: the first lmax = 2^k (k >= 1) that fails, all probes at once
  int32_t lmax = 1 << 19;
  {
    int32_t v[18];
#pragma unroll
    for (int k = 1; k <= 18; ++k) v[k - 1] = delta(i + (1 << k) * d);
#pragma unroll
    for (int k = 18; k >= 1; --k)
      if (!(v[k - 1] > dmin)) lmax = 1 << k;
  }
  const int32_t l = last_pass(lmax >= 4 ? lmax / 2 : 0, lmax - 1, d, dmin);
  const int32_t j = i + l * d;
  const int32_t dnode = delta(j);
  const int32_t s = last_pass(0, l - 1, d, dnode);
  const int32_t split = i + s * d + (d < 0 ? d : 0);
  const int32_t lo = (i < j) ? i : j;
  const int32_t hi = (i < j) ? j : i;
  range[3 * i] = lo;
  range[3 * i + 1] = hi;
  range[3 * i + 2] = split;
  if (lo == split) {
    parents[split + inner] = i;
    lchild[i] = split + inner;
  } else {
    parents[split] = i;
    lchild[i] = split;
  }
  if (hi == split + 1) {
    parents[split + inner + 1] = i;
    rchild[i] = split + inner + 1;
  } else {
    parents[split + 1] = i;
    rchild[i] = split + 1;
  }
  if (i == 0) parents[0] = -1;
  lo_out = lo;
  hi_out = hi;
  split_out = split;
}

// The same launch clears the refit's counters (zero_a[0, n_a), zero_b[0, n_b)).
__global__ __launch_bounds__(kBlock) void k_karras(const uint32_t* __restrict__ mc, int32_t inner,
                                                    int32_t* __restrict__ lchild,
                                                    int32_t* __restrict__ rchild,
                                                    int32_t* __restrict__ parents,
                                                    int32_t* __restrict__ range,
                                                    int32_t* __restrict__ zero_a, int32_t n_a,
                                                    int32_t* __restrict__ zero_b, int32_t n_b) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n_a) zero_a[i] = 0;
  if (i < n_b) zero_b[i] = 0;
  if (i >= inner) return;
  karras_node(mc, inner, i, lchild, rchild, parents, range);
}

// ---- box unions over leaf ranges (small builds) ------------------------------
// BoundingBox::addBox is associative over boxes without a signed-zero pair
// (box_signed_zero_pair), so the box of a run of leaves can be put together
// from any split of the run into consecutive pieces, combined left to right
// as the SEQ refit combines a node's children: the right part's box, then
// addBox(the left part's box) (the second arrival at a node in leaf order
// comes from its right child, build_radix_tree.hpp:536-571). The pieces come
// from a segment tree over the leaves in tree order: P = the power of two at
// or above n, node x in [1, P) the union of leaves [(x << l) - P, ((x + 1) << l) - P)
// at level l, leaf k at P + k; leaves past n are absent.
struct SegBox {
  double b[6];
  bool empty;
};

__device__ __forceinline__ void seg_combine(SegBox& left, const SegBox& right) {  // left := left ++ right
  if (right.empty) return;
  if (left.empty) {
    left = right;
    return;
  }
  double r[6];
  for (int k = 0; k < 6; ++k) r[k] = right.b[k];
  box_add_box(r, left.b);
  for (int k = 0; k < 6; ++k) left.b[k] = r[k];
}

__device__ __forceinline__ SegBox seg_shfl_down(const SegBox& v, int off) {
  SegBox o;
  for (int k = 0; k < 6; ++k) o.b[k] = __shfl_down(v.b[k], off);
  o.empty = __shfl_down((int)v.empty, off) != 0;
  return o;
}

constexpr int kWaveLanes = 64;

// propagate_aabbs, build_radix_tree.hpp:505-576: every leaf thread climbs
// towards the root; at each node the first thread to arrive stops and the
// second unions the two child boxes and goes on. Here the climb is split by
// leaf spans so that every hand-off stays inside one workgroup (on gfx950 an
// agent-scope release/acquire writes back and invalidates the XCD's L2, which
// per node would cost far more than the refit itself). Inner node i of a
// Karras tree covers a contiguous leaf range and its index is that range's
// low or high end, so the nodes whose range lies inside a span of leaves form
// subtrees of that span.
//   k_refit_chunks  spans of kBlock leaves: links, counters and boxes in LDS
//   k_refit_span    spans of kRefitSpan leaves, one workgroup each (counters
//                   in LDS), then one workgroup for the rest of the tree
//                   (global counters); boxes in global memory, workgroup-scope
//                   hand-off
// A thread that reaches a node whose range leaves its span records the child
// it comes from (an exit) for the next tier and stops. No thread waits on
// another, and every thread stops after at most depth-many steps.
constexpr int kRefitThreads = 1024;

// A node's box from the climbing thread's box (acc) and the other child's
// (ob), in the SEQ refit's order whichever child the thread came from: the
// right child's box, then addBox(the left child's box). acc := the result.
__device__ __forceinline__ void refit_union(double acc[6], double ob[6], bool ob_is_right) {
  if (ob_is_right) {
    box_add_box(ob, acc);
    for (int k = 0; k < 6; ++k) acc[k] = ob[k];
  } else {
    box_add_box(acc, ob);
  }
}
constexpr int64_t kRefitSpan = 16384;  // a multiple of kBlock

constexpr int kChunksPerSpan = (int)(kRefitSpan / kBlock);
static_assert(kChunksPerSpan == 64, "one wave scans the exit counts of a span's chunks");

// The chunk's own inner nodes (indices [base, base + kBlock)) and leaves are
// staged in LDS, and the climb runs in LDS alone: the boxes it makes are
// written out afterwards in one coalesced pass, and the chunk's exits go to
// its own kBlock slots of `exits`, with their number in n_exits[chunk] (a
// chunk's exits root disjoint sets of its leaves, so there are at most kBlock).
__global__ __launch_bounds__(kBlock) void k_refit_chunks(const double* __restrict__ leaf_boxes,
                                                          int32_t inner,
                                                          const int32_t* __restrict__ lchild,
                                                          const int32_t* __restrict__ rchild,
                                                          const int32_t* __restrict__ parents,
                                                          const int32_t* __restrict__ range,
                                                          double* __restrict__ inner_boxes,
                                                          int32_t* __restrict__ exits,
                                                          int32_t* __restrict__ n_exits) {
  __shared__ int32_t s_cnt[kBlock];
  __shared__ int32_t s_lo[kBlock], s_hi[kBlock], s_lc[kBlock], s_rc[kBlock], s_par[kBlock];
  __shared__ int32_t s_ex[kBlock];
  __shared__ int32_t s_nex;
  __shared__ double s_box[6][kBlock];
  __shared__ double s_leaf[6][kBlock];
  const int32_t t = (int32_t)threadIdx.x;
  const int32_t base = (int32_t)(blockIdx.x * kBlock);
  const int32_t top = base + kBlock - 1;  // the chunk is leaves [base, top]
  const int32_t i = base + t;             // leaf i and inner node i
  s_cnt[t] = 0;
  if (t == 0) s_nex = 0;
  if (i < inner) {
    s_lo[t] = range[3 * i];
    s_hi[t] = range[3 * i + 1];
    s_lc[t] = lchild[i];
    s_rc[t] = rchild[i];
    s_par[t] = parents[i];
  }
  double acc[6];
  int32_t cur = -1;
  if (i <= inner) {
    load_box(leaf_boxes + 6 * (int64_t)i, acc);
    for (int k = 0; k < 6; ++k) s_leaf[k][t] = acc[k];
    cur = parents[inner + i];
  }
  __syncthreads();
  if (i <= inner) {
    int32_t last = inner + i;
    while (cur != -1) {
      // a node outside [base, top] has its range's low or high end there
      const int32_t c = cur - base;
      if (c < 0 || c >= kBlock || s_lo[c] < base || s_hi[c] > top) {
        s_ex[atomicAdd(&s_nex, 1)] = last;
        break;
      }
      if (__hip_atomic_fetch_add(&s_cnt[c], 1, __ATOMIC_ACQ_REL,
                                 __HIP_MEMORY_SCOPE_WORKGROUP) == 0)
        break;
      const int32_t lc = s_lc[c], rc = s_rc[c];
      const int32_t other = (lc == last) ? rc : lc;  // a chunk node or a chunk leaf
      double ob[6];
      if (other >= inner)
        for (int k = 0; k < 6; ++k) ob[k] = s_leaf[k][other - inner - base];
      else
        for (int k = 0; k < 6; ++k) ob[k] = s_box[k][other - base];
      refit_union(acc, ob, other == rc);
      for (int k = 0; k < 6; ++k) s_box[k][c] = acc[k];
      last = cur;
      cur = s_par[c];
    }
  }
  __syncthreads();
  if (i < inner && s_lo[t] >= base && s_hi[t] <= top) {  // one of the chunk's own nodes
    double b[6];
    for (int k = 0; k < 6; ++k) b[k] = s_box[k][t];
    store_box(inner_boxes + 6 * (int64_t)i, b);
  }
  const int32_t ne = s_nex;
  for (int32_t j = t; j < ne; j += kBlock) exits[(int64_t)blockIdx.x * kBlock + j] = s_ex[j];
  if (t == 0) n_exits[blockIdx.x] = ne;
}

// Above the chunks. FROM_CHUNKS: workgroup g climbs from the exits of chunks
// [g * kChunksPerSpan, (g + 1) * kChunksPerSpan) (kBlock slots per chunk,
// counts per chunk, numbered across the span by a scan of the counts) through the nodes whose
// range lies in leaves [g * kRefitSpan, (g + 1) * kRefitSpan - 1], with the
// arrival counters in LDS (a node of the span has its index in the span), and
// records where it leaves the span in exits_out (one list, count *n_out).
// Otherwise one workgroup climbs from that list to the root with global
// counters[node]. Boxes below the exits were stored by the previous kernel;
// within a kernel the hand-off is between waves of one workgroup. A node's
// links and range are loaded one level ahead, while the hand-off at the node
// below is in flight.
struct RefitTopo {
  int32_t lo, hi, lc, rc, par;
};

__device__ __forceinline__ RefitTopo refit_topo(const int32_t* __restrict__ lchild,
                                                const int32_t* __restrict__ rchild,
                                                const int32_t* __restrict__ parents,
                                                const int32_t* __restrict__ range, int32_t c) {
  const int32_t k = c < 0 ? 0 : c;  // the root's parent (-1): a harmless read
  return RefitTopo{range[3 * k], range[3 * k + 1], lchild[k], rchild[k], parents[k]};
}

template <bool FROM_CHUNKS>
__global__ __launch_bounds__(kRefitThreads) void k_refit_span(
    const double* __restrict__ leaf_boxes, int32_t inner, const int32_t* __restrict__ lchild,
    const int32_t* __restrict__ rchild, const int32_t* __restrict__ parents,
    const int32_t* __restrict__ range, const int32_t* __restrict__ exits_in,
    const int32_t* __restrict__ n_in, int64_t nchunks, int32_t* __restrict__ exits_out,
    int32_t* __restrict__ n_out, int32_t* counters, double* inner_boxes) {
  __shared__ int32_t s_cnt[FROM_CHUNKS ? kRefitSpan : 1];
  __shared__ int32_t s_off[kChunksPerSpan + 1];  // exits of the span's chunks, scanned
  const int32_t t = (int32_t)threadIdx.x;
  int64_t lo = 0, hi = (int64_t)inner;  // leaves of the span
  int32_t ne;
  const int64_t c0 = (int64_t)blockIdx.x * kChunksPerSpan;
  if constexpr (FROM_CHUNKS) {
    lo = (int64_t)blockIdx.x * kRefitSpan;
    hi = lo + kRefitSpan - 1;
    for (int32_t j = t; j < (int32_t)kRefitSpan; j += kRefitThreads) s_cnt[j] = 0;
    if (t < kChunksPerSpan) {  // one wave: inclusive scan of the chunks' exit counts
      int32_t v = (c0 + t < nchunks) ? n_in[c0 + t] : 0;
      for (int off = 1; off < kChunksPerSpan; off <<= 1) {
        const int32_t y = __shfl_up(v, off, kChunksPerSpan);
        if (t >= off) v += y;
      }
      s_off[t + 1] = v;
      if (t == 0) s_off[0] = 0;
    }
    __syncthreads();
    ne = s_off[kChunksPerSpan];
  } else {
    ne = *n_in;
  }
  for (int32_t e = t; e < ne; e += kRefitThreads) {
    int32_t last;
    if constexpr (FROM_CHUNKS) {  // the chunk k with s_off[k] <= e < s_off[k + 1]
      int32_t a = 0, b = kChunksPerSpan;
      while (b - a > 1) {
        const int32_t m = (a + b) >> 1;
        if (s_off[m] <= e)
          a = m;
        else
          b = m;
      }
      last = exits_in[(c0 + a) * kBlock + (e - s_off[a])];
    } else {
      last = exits_in[e];
    }
    double acc[6];
    if (last >= inner)
      load_box(leaf_boxes + 6 * (int64_t)(last - inner), acc);
    else
      load_box(inner_boxes + 6 * (int64_t)last, acc);
    int32_t cur = parents[last];
    RefitTopo tp = refit_topo(lchild, rchild, parents, range, cur);
    while (cur != -1) {
      if (tp.lo < lo || tp.hi > hi) {
        if (exits_out) exits_out[atomicAdd(n_out, 1)] = last;  // none above the spans
        break;
      }
      const int32_t nxt = tp.par;
      const RefitTopo tn = refit_topo(lchild, rchild, parents, range, nxt);
      int32_t* ctr;
      if constexpr (FROM_CHUNKS)
        ctr = &s_cnt[cur - lo];
      else
        ctr = &counters[cur];
      if (__hip_atomic_fetch_add(ctr, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0)
        break;
      const int32_t other = (tp.lc == last) ? tp.rc : tp.lc;
      double ob[6];
      if (other >= inner)
        load_box(leaf_boxes + 6 * (int64_t)(other - inner), ob);
      else
        load_box(inner_boxes + 6 * (int64_t)other, ob);
      refit_union(acc, ob, other == tp.rc);
      store_box(inner_boxes + 6 * (int64_t)cur, acc);
      last = cur;
      cur = nxt;
      tp = tn;
    }
  }
}

__device__ __forceinline__ float f_down(double x) { return __double2float_rd(x); }
__device__ __forceinline__ float f_up(double x) { return __double2float_ru(x); }

// emit_bvh_parents, LinearBVH.hpp:225-264, into 128-byte nodes (child links
// are node indices, not doubled offsets) and, for mixed precision, 64-byte
// FP32 nodes whose min corners round down and max corners round up. lc / rc:
// inner node index, or -(leaf + 1).
__device__ __forceinline__ void emit_node(int32_t i, int32_t lc, int32_t rc, const double lb[6],
                                          const double rb[6], int32_t lo, int32_t hi,
                                          int32_t mid, Node64* __restrict__ nodes,
                                          Node32* __restrict__ nodes32,
                                          int4* __restrict__ ranges32) {
  Node64 nd;
  for (int k = 0; k < 6; ++k) {
    nd.l[k] = lb[k];
    nd.r[k] = rb[k];
  }
  nd.lc = lc;
  nd.rc = rc;
  nd.flags = (box_valid(nd.l) ? 1 : 0) | (box_valid(nd.r) ? 2 : 0);
  nd.lo = lo;
  nd.hi = hi;
  nd.mid = mid;
  for (int k = 0; k < 2; ++k) nd.pad[k] = 0;
  nodes[i] = nd;
  if (nodes32) {
    Node32 h;
    for (int k = 0; k < 3; ++k) {
      h.l[k] = f_down(nd.l[k]);
      h.l[3 + k] = f_up(nd.l[3 + k]);
      h.r[k] = f_down(nd.r[k]);
      h.r[3 + k] = f_up(nd.r[3 + k]);
    }
    h.lc = lc;
    h.rc = rc;
    h.flags = (box_valid_f(h.l) ? 1 : 0) | (box_valid_f(h.r) ? 2 : 0);
    h.pad = 0;
    nodes32[i] = h;
    ranges32[i] = make_int4(lo, hi, mid, 0);
  }
}

__global__ __launch_bounds__(kBlock) void k_emit(const double* __restrict__ leaf_boxes,
                                                  const double* __restrict__ inner_boxes,
                                                  int32_t inner,
                                                  const int32_t* __restrict__ lchild,
                                                  const int32_t* __restrict__ rchild,
                                                  const int32_t* __restrict__ range,
                                                  Node64* __restrict__ nodes,
                                                  Node32* __restrict__ nodes32,
                                                  int4* __restrict__ ranges32,
                                                  int32_t* __restrict__ status) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i == 0) *status = 0;  // the tree's error flag starts clear
  if (i >= inner) return;
  double lb[6], rb[6];
  int32_t lc = lchild[i], rc = rchild[i];
  if (lc >= inner) {
    load_box(leaf_boxes + 6 * (int64_t)(lc - inner), lb);
    lc = -(lc - inner + 1);
  } else {
    load_box(inner_boxes + 6 * (int64_t)lc, lb);
  }
  if (rc >= inner) {
    load_box(leaf_boxes + 6 * (int64_t)(rc - inner), rb);
    rc = -(rc - inner + 1);
  } else {
    load_box(inner_boxes + 6 * (int64_t)rc, rb);
  }
  emit_node(i, lc, rc, lb, rb, range[3 * i], range[3 * i + 1], range[3 * i + 2], nodes, nodes32,
            ranges32);
}

constexpr int kSegLeaves = 512;  // leaves per workgroup of k_seg_local
constexpr int kSegLog = 9;

// Small builds, one 1,024-thread workgroup per 512 leaves and inner nodes
// (base = 512 * workgroup). Its first 512 threads find the hierarchy's nodes
// base .. base + 511 (karras_node_wide) while the other 512 put together the
// segment tree of its leaves in LDS (local node u in [1, 512) at level l covers leaves
// base + [(u << l) - 512, ((u + 1) << l) - 512), leaf base + k sits at
// 512 + k; levels 1-6 within each wave by shuffles, 7-9 in the first leaf
// wave) and stores its nodes with a leaf at their global places (levels
// 1 .. min(9, log2 P)). Then every inner node whose leaf range lies in the
// workgroup's leaves is emitted with child boxes taken from LDS; the others
// are listed for k_emit_seg, kSegLeaves slots per workgroup
// (deferred[512 * workgroup + j], their number in ndeferred[workgroup]).
__global__ __launch_bounds__(2 * kSegLeaves) void k_seg_local(
    const uint32_t* __restrict__ mc, int32_t inner, int32_t* __restrict__ lchild,
    int32_t* __restrict__ rchild, int32_t* __restrict__ parents, int32_t* __restrict__ range,
    const double* __restrict__ leaf_boxes, int64_t n, int64_t P, int levels,
    double* __restrict__ seg, Node64* __restrict__ nodes, Node32* __restrict__ nodes32,
    int4* __restrict__ ranges32, int32_t* __restrict__ status, int32_t* __restrict__ deferred,
    int32_t* __restrict__ ndeferred) {
  __shared__ double s_box[6][2 * kSegLeaves];
  __shared__ int32_t s_rng[3][kSegLeaves];
  __shared__ int32_t s_ndef;
  const int t = threadIdx.x;
  const int lane = t % kWaveLanes, wid = t / kWaveLanes;
  const int64_t base = (int64_t)blockIdx.x * kSegLeaves;
  if (base + t == 0) *status = 0;  // the tree's error flag starts clear
  if (t == 0) s_ndef = 0;
  // node u of local level l: stored at s_box[.][u] and, with a leaf, globally
  auto put = [&](const SegBox& x, int l, int u, int64_t first_leaf) {
    for (int c = 0; c < 6; ++c) s_box[c][u] = x.b[c];
    if (l <= levels && !x.empty) store_box(seg + 6 * ((P >> l) + (first_leaf >> l)), x.b);
  };
  if (t < kSegLeaves) {  // the first 8 waves: the hierarchy's node base + t
    const int64_t kn = base + t;
    int32_t nlo = 0, nhi = 0, nsp = 0;
    if (kn < inner)
      karras_node_wide(mc, inner, (int32_t)kn, lchild, rchild, parents, range, nlo, nhi, nsp);
    s_rng[0][t] = nlo;
    s_rng[1][t] = nhi;
    s_rng[2][t] = nsp;
  } else {  // the other 8: leaf base + t - 512 and the tree's levels 1-6
    const int tl = t - kSegLeaves;
    const int64_t kl = base + tl;
    SegBox v;
    v.empty = kl >= n;
    if (!v.empty)
      load_box(leaf_boxes + 6 * kl, v.b);
    else
      invalid_box(v.b);
    for (int c = 0; c < 6; ++c) s_box[c][kSegLeaves + tl] = v.b[c];
    for (int l = 1; l <= 6; ++l) {
      const SegBox o = seg_shfl_down(v, 1 << (l - 1));
      if ((lane & ((1 << l) - 1)) == 0) {
        seg_combine(v, o);
        put(v, l, (kSegLeaves >> l) + (tl >> l), kl);
      }
    }
  }
  __syncthreads();  // the leaf waves' level-6 nodes, local 8 .. 15 (leaf wave w: 8 + w)
  if (wid == kSegLeaves / kWaveLanes) {  // the first leaf wave: levels 7-9
    const int nw = kSegLeaves / kWaveLanes;  // 8
    SegBox u;
    for (int c = 0; c < 6; ++c) u.b[c] = s_box[c][nw + (lane < nw ? lane : 0)];
    u.empty = lane >= nw || base + (int64_t)lane * kWaveLanes >= n;
    for (int l = 7; l <= kSegLog; ++l) {
      const SegBox o = seg_shfl_down(u, 1 << (l - 7));
      if ((lane & ((1 << (l - 6)) - 1)) == 0 && lane < nw) {
        seg_combine(u, o);
        put(u, l, (kSegLeaves >> l) + (lane >> (l - 6)), base + (int64_t)lane * kWaveLanes);
      }
    }
  }
  // node pairs: threads 2j and 2j + 1 take inner node base + j, its left and
  // right child
  const int64_t k = base + (t >> 1);
  const int side = t & 1;
  const int32_t lo = s_rng[0][t >> 1], hi = s_rng[1][t >> 1], split = s_rng[2][t >> 1];
  const bool mine = k < inner && lo >= base && hi < base + kSegLeaves;
  if (side == 0 && k < inner && !mine) deferred[base + atomicAdd(&s_ndef, 1)] = (int32_t)k;
  __syncthreads();
  if (t == 0) ndeferred[blockIdx.x] = s_ndef;
  if (!mine) return;  // both threads of a pair leave together
  // union of local leaves [a, b] from LDS, pieces combined in leaf order
  auto local_box = [&](int32_t a, int32_t b, double out[6]) {
    int32_t l = a + kSegLeaves, r = b + kSegLeaves + 1;
    SegBox L, R;
    L.empty = R.empty = true;
    while (l < r) {
      if (l & 1) {
        SegBox o;
        for (int c = 0; c < 6; ++c) o.b[c] = s_box[c][l];
        o.empty = false;
        seg_combine(L, o);
        ++l;
      }
      if (r & 1) {
        --r;
        SegBox o;
        for (int c = 0; c < 6; ++c) o.b[c] = s_box[c][r];
        o.empty = false;
        seg_combine(o, R);
        R = o;
      }
      l >>= 1;
      r >>= 1;
    }
    seg_combine(L, R);
    for (int c = 0; c < 6; ++c) out[c] = L.b[c];
  };
  const int32_t b0 = (int32_t)base;
  // this thread's child: [lo, split] (side 0) or [split + 1, hi] (side 1)
  const int32_t ca = side ? split + 1 : lo, cb = side ? hi : split;
  double bx[6];
  if (ca == cb) {
    for (int c = 0; c < 6; ++c) bx[c] = s_box[c][kSegLeaves + ca - b0];
  } else {
    local_box(ca - b0, cb - b0, bx);
  }
  double rb[6];
  for (int c = 0; c < 6; ++c) rb[c] = __shfl_xor(bx[c], 1);
  if (side == 0) {
    const int32_t lc = (lo == split) ? -(split + 1) : split;
    const int32_t rc = (hi == split + 1) ? -(split + 2) : split + 1;
    emit_node((int32_t)k, lc, rc, bx, rb, lo, hi, split, nodes, nodes32, ranges32);
  }
}

// Order-keeping union over each half of the wave: lanes 0 and 32 end with
// the union of their half's lanes in lane order.
__device__ __forceinline__ SegBox half_wave_union(SegBox v) {
  const int hl = threadIdx.x % (kWaveLanes / 2);
  for (int off = 1; off < kWaveLanes / 2; off <<= 1) {
    const SegBox o = seg_shfl_down(v, off);
    if ((hl & (2 * off - 1)) == 0) seg_combine(v, o);
  }
  return v;
}

// The deferred nodes (their leaf range reaches outside one workgroup of
// k_seg_local), one wave each (the waves split evenly over the workgroups'
// lists), its two halves on the two children at once. A child's leaf range
// [a, b] is the pieces the bottom-up walk takes below level 9 (left side
// ascending, right side descending; at most 9 each) and the run of level-9
// nodes between them. Lane j of a half takes the j-th left piece, the j-th
// right piece from the right end and up to 16 consecutive level-9 nodes of
// the run (loaded 8 at a time); three order-keeping unions over the half
// then give left ++ run ++ right.
__global__ __launch_bounds__(kBlock) void k_emit_seg(
    const double* __restrict__ leaf_boxes, const double* __restrict__ seg, int64_t P,
    const int32_t* __restrict__ range, const int32_t* __restrict__ deferred,
    const int32_t* __restrict__ ndeferred, int64_t nlists, Node64* __restrict__ nodes,
    Node32* __restrict__ nodes32, int4* __restrict__ ranges32) {
  constexpr int kHalf = kWaveLanes / 2;
  const int lane = threadIdx.x % kWaveLanes;
  const int half = lane / kHalf, hl = lane % kHalf;
  const int64_t wave = (int64_t)blockIdx.x * (kBlock / kWaveLanes) + threadIdx.x / kWaveLanes;
  // the waves split evenly over the lists: list g, entries q0, q0 + per, ...
  const int64_t per = ((int64_t)gridDim.x * (kBlock / kWaveLanes)) / nlists;
  if (wave >= per * nlists) return;
  const int64_t g = wave / per, q0 = wave % per;
  const int32_t nd = ndeferred[g];
  auto piece = [&](int64_t x, SegBox& o) {
    load_box(x >= P ? leaf_boxes + 6 * (x - P) : seg + 6 * x, o.b);
    o.empty = false;
  };
  // union of leaves [a, b] over this half of the wave (the result in its lane 0)
  auto range_box = [&](int64_t a, int64_t b, double out[6]) {
    int64_t l = a + P, r = b + 1 + P;
    int nr = 0;
    for (int lev = 0; lev < kSegLog && l < r; ++lev) {
      if (l & 1) ++l;
      if (r & 1) ++nr, --r;
      l >>= 1;
      r >>= 1;
    }
    const int64_t m0 = l, m1 = l < r ? r : l;  // the level-9 run (when the walk got there)
    int64_t xl = -1, xr = -1;
    l = a + P;
    r = b + 1 + P;
    int cl = 0, cr = 0;
    for (int lev = 0; lev < kSegLog && l < r; ++lev) {
      if (l & 1) {
        if (cl == hl) xl = l;
        ++cl;
        ++l;
      }
      if (r & 1) {
        --r;
        if (nr - 1 - cr == hl) xr = r;
        ++cr;
      }
      l >>= 1;
      r >>= 1;
    }
    const int64_t per_lane = (m1 - m0 + kHalf - 1) / kHalf;  // at most 16
    const int64_t j0 = m0 + per_lane * hl, j1 = (j0 + per_lane < m1) ? j0 + per_lane : m1;
    SegBox L, M, R, mb[8];
    L.empty = xl < 0;
    R.empty = xr < 0;
    if (xl >= 0) piece(xl, L);
    if (xr >= 0) piece(xr, R);
    M.empty = true;
    for (int64_t jb = j0; jb < j1; jb += 8) {
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (jb + u < j1) piece(jb + u, mb[u]);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (jb + u < j1) seg_combine(M, mb[u]);
    }
    L = half_wave_union(L);
    M = half_wave_union(M);
    R = half_wave_union(R);
    seg_combine(L, M);
    seg_combine(L, R);
    for (int c = 0; c < 6; ++c) out[c] = L.b[c];
  };
  for (int64_t q = q0; q < nd; q += per) {
    const int32_t i = deferred[g * kSegLeaves + q];
    const int32_t lo = range[3 * i], hi = range[3 * i + 1], split = range[3 * i + 2];
    // half 0: the left child [lo, split]; half 1: the right child [split + 1, hi]
    const int32_t ca = half ? split + 1 : lo, cb = half ? hi : split;
    double bx[6];
    if (ca == cb)
      load_box(leaf_boxes + 6 * (int64_t)ca, bx);
    else
      range_box(ca, cb, bx);
    double rb[6];
    for (int c = 0; c < 6; ++c) rb[c] = __shfl(bx[c], kHalf);
    if (lane == 0) {
      const int32_t lc = (lo == split) ? -(split + 1) : split;
      const int32_t rc = (hi == split + 1) ? -(split + 2) : split + 1;
      emit_node(i, lc, rc, bx, rb, lo, hi, split, nodes, nodes32, ranges32);
    }
  }
}

// Reference layout (LinearBVH.hpp:170-187,257-262): child boxes of inner
// node i at 2i and 2i+1, child links doubled for inner children.
__global__ __launch_bounds__(kBlock) void k_export(const Node64* __restrict__ nodes, int32_t inner,
                                                    int dims, double* __restrict__ boxes,
                                                    int32_t* __restrict__ children) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= inner) return;
  const Node64 nd = nodes[i];
  if (boxes && dims == 2) {  // BoundingBox<double, 2>: min x y, max x y
    const int src[4] = {0, 1, 3, 4};
    for (int k = 0; k < 4; ++k) {
      boxes[8 * (int64_t)i + k] = nd.l[src[k]];
      boxes[8 * (int64_t)i + 4 + k] = nd.r[src[k]];
    }
  } else if (boxes) {
    for (int k = 0; k < 6; ++k) {
      boxes[12 * (int64_t)i + k] = nd.l[k];
      boxes[12 * (int64_t)i + 6 + k] = nd.r[k];
    }
  }
  if (children) {
    children[2 * (int64_t)i] = nd.lc >= 0 ? 2 * nd.lc : nd.lc;
    children[2 * (int64_t)i + 1] = nd.rc >= 0 ? 2 * nd.rc : nd.rc;
  }
}

// 2-D boxes (min x y, max x y) as 3-D boxes with z range [0, 0]: every test,
// union, scaling and bound of the build and the searches then gives the 2-D
// result, since the z terms are exact zeros.
__global__ __launch_bounds__(kBlock) void k_widen_boxes2d(const double* __restrict__ in,
                                                          int64_t n, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* b = in + 4 * i;
  double* o = out + 6 * i;
  o[0] = b[0];
  o[1] = b[1];
  o[2] = 0.0;
  o[3] = b[2];
  o[4] = b[3];
  o[5] = 0.0;
}

}  // namespace

int widen_boxes2d(const double* boxes2d, int64_t n, double* out, hipStream_t s) {
  if (n <= 0) return AXSQ_OK;
  k_widen_boxes2d<<<grid_for(n), kBlock, 0, s>>>(boxes2d, n, out);
  AXSQ_CHECK_HIP(hipGetLastError());
  return AXSQ_OK;
}

int bvh_export_launch(const axsq_bvh* b, double* inner_nodes, int32_t* children,
                      int32_t* leaf_nodes, uint32_t* mcodes, hipStream_t s) {
  const int32_t inner = (int32_t)(b->n - 1);
  if (inner_nodes || children) {
    k_export<<<grid_for(inner), kBlock, 0, s>>>(b->nodes, inner, b->dims, inner_nodes, children);
    AXSQ_CHECK_HIP(hipGetLastError());
  }
  if (leaf_nodes)
    AXSQ_CHECK_HIP(hipMemcpyAsync(leaf_nodes, b->leaf_ids, sizeof(int32_t) * b->n,
                                  hipMemcpyDeviceToDevice, s));
  if (mcodes)
    AXSQ_CHECK_HIP(hipMemcpyAsync(mcodes, b->mcodes, sizeof(uint32_t) * b->n,
                                  hipMemcpyDeviceToDevice, s));
  return AXSQ_OK;
}

void bvh_release(axsq_bvh* b) {
  if (!b) return;
  if (b->build_graph) hipGraphExecDestroy(b->build_graph);
  b->build_graph = nullptr;
  b->build_graph_key.clear();
  if (b->host_params) hipHostFree(b->host_params);
  b->host_params = b->host_params_dev = nullptr;
  hipFree(b->nodes);
  hipFree(b->nodes32);
  hipFree(b->ranges32);
  hipFree(b->leaf_boxes);
  hipFree(b->leaf_ids);
  hipFree(b->mcodes);
  hipFree(b->status);
  hipFree(b->find_visits);
  b->nodes = nullptr;
  b->nodes32 = nullptr;
  b->ranges32 = nullptr;
  b->leaf_boxes = nullptr;
  b->leaf_ids = nullptr;
  b->mcodes = nullptr;
  b->status = nullptr;
  b->find_visits = nullptr;
}

template <class T>
static int dmalloc(T** p, size_t count) {
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (count ? count : 1));
  if (e != hipSuccess) {
    set_error("hipMalloc of %zu bytes failed: %s", sizeof(T) * count, hipGetErrorString(e));
    return AXSQ_ERR_NOMEM;
  }
  return AXSQ_OK;
}

// A stream of the current device for graph capture (created once per
// device; capture needs a stream other than the null stream).
static hipStream_t capture_stream() {
  static hipStream_t streams[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  if (!streams[dev] && hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking) != hipSuccess)
    streams[dev] = nullptr;
  return streams[dev];
}

int bvh_build_into(axsq_bvh* b, const double* boxes, int64_t n_in, double scale, int fp_mode,
                   hipStream_t s, int dims, bool f32, const float* boxes_f32) {
  if (n_in < 0 || n_in > 0x3fffffff) {
    set_error("axsq_bvh_build: box count %lld out of range", (long long)n_in);
    return AXSQ_ERR_INVALID_ARG;
  }
  if (n_in > 0 && (f32 ? !boxes_f32 : !boxes)) {
    set_error("axsq_bvh_build: boxes is NULL");
    return AXSQ_ERR_INVALID_ARG;
  }
  const int64_t n = (n_in <= 1) ? 2 : n_in;
  const int32_t inner = (int32_t)(n - 1);
  Workspace& ws = workspace();

  // (re)allocate the persistent arrays when the size changes
  if (b->n != n || (fp_mode == AXSQ_FP32_TRAVERSAL && !b->nodes32) || !b->nodes) {
    bvh_release(b);
    AXSQ_TRY(dmalloc(&b->nodes, (size_t)inner));
    if (fp_mode == AXSQ_FP32_TRAVERSAL) {
      AXSQ_TRY(dmalloc(&b->nodes32, (size_t)inner));
      AXSQ_TRY(dmalloc(&b->ranges32, (size_t)inner));
    }
    AXSQ_TRY(dmalloc(&b->leaf_boxes, 6 * (size_t)n));
    AXSQ_TRY(dmalloc(&b->leaf_ids, (size_t)n));
    AXSQ_TRY(dmalloc(&b->mcodes, (size_t)n));
    AXSQ_TRY(dmalloc(&b->status, 1));
    AXSQ_TRY(dmalloc(&b->find_visits, 1));
  }
  b->n = n;
  b->fp_mode = fp_mode;
  b->scale = scale;
  b->dims = dims;
  b->f32 = f32;

  // scratch: scaled boxes (6n doubles), keys/vals x2, tree arrays, partials
  const int64_t nparts = grid_for(n);
  AXSQ_TRY(ws.k0.ensure(sizeof(uint32_t) * n));
  AXSQ_TRY(ws.v0.ensure(sizeof(uint32_t) * n));
  AXSQ_TRY(ws.k1.ensure(sizeof(uint32_t) * n));
  AXSQ_TRY(ws.v1.ensure(sizeof(uint32_t) * n));
  // inputs of up to kSmallSortMax boxes take the segment-tree path
  const bool small = n <= kSmallSortMax;
  int64_t P = 2;
  int levels = 1;
  while (P < n) {
    P *= 2;
    ++levels;
  }
  // flag buffer holds: scaled boxes | partials | params | lchild | rchild |
  // parents | counters | leaf ranges | inner boxes | refit exits | segment tree
  const size_t off_part = sizeof(double) * 6 * n;
  const size_t off_prm = off_part + sizeof(double) * kPartStride * nparts;
  const size_t off_l = (off_prm + sizeof(BuildParams) + 255) & ~size_t(255);
  const size_t off_r = off_l + sizeof(int32_t) * inner;
  const size_t off_p = off_r + sizeof(int32_t) * inner;
  const size_t off_c = off_p + sizeof(int32_t) * (2 * n - 1);
  const size_t off_rg = off_c + sizeof(int32_t) * inner;
  const size_t off_ib = (off_rg + sizeof(int32_t) * 3 * inner + 255) & ~size_t(255);
  // refit exits: per-chunk counts of the first tier, one count for the
  // second, then the exit lists (kBlock slots per chunk; one list)
  const int64_t nchunks = (n + kBlock - 1) / kBlock;
  const int64_t nspans = (n + kRefitSpan - 1) / kRefitSpan;
  const size_t off_ex = off_ib + sizeof(double) * 6 * inner;
  const size_t off_seg =
      (off_ex + sizeof(int32_t) * (nchunks + 1 + nchunks * kBlock + n) + 255) & ~size_t(255);
  const int64_t nseg = (n + kSegLeaves - 1) / kSegLeaves;  // workgroups of k_seg_local
  const size_t off_def = off_seg + (small ? sizeof(double) * 6 * (size_t)P : 0);
  const size_t total = off_def + (small ? sizeof(int32_t) * (size_t)nseg * (kSegLeaves + 1) : 0);
  AXSQ_TRY(ws.flag.ensure(total));
  char* base = ws.flag.as<char>();
  double* scaled = reinterpret_cast<double*>(base);
  double* partials = reinterpret_cast<double*>(base + off_part);
  BuildParams* prm = reinterpret_cast<BuildParams*>(base + off_prm);
  int32_t* lch = reinterpret_cast<int32_t*>(base + off_l);
  int32_t* rch = reinterpret_cast<int32_t*>(base + off_r);
  int32_t* par = reinterpret_cast<int32_t*>(base + off_p);
  int32_t* cnt = reinterpret_cast<int32_t*>(base + off_c);
  int32_t* rng = reinterpret_cast<int32_t*>(base + off_rg);
  double* ib = reinterpret_cast<double*>(base + off_ib);
  int32_t* n_ex0 = reinterpret_cast<int32_t*>(base + off_ex);
  int32_t* n_ex1 = n_ex0 + nchunks;
  int32_t* ex0 = n_ex1 + 1;
  int32_t* ex1 = ex0 + nchunks * kBlock;
  double* seg = reinterpret_cast<double*>(base + off_seg);
  int32_t* deferred = reinterpret_cast<int32_t*>(base + off_def);  // per k_seg_local workgroup:
  int32_t* ndef = deferred + nseg * kSegLeaves;                      // a list and its length
  Node32* n32 = fp_mode == AXSQ_FP32_TRAVERSAL ? b->nodes32 : nullptr;

  // the exact climb: refit tiers, then the nodes (counters cleared by the caller)
  auto climb_and_emit = [&]() -> int {
    k_refit_chunks<<<grid_for(n), kBlock, 0, s>>>(b->leaf_boxes, inner, lch, rch, par, rng, ib,
                                                  ex0, n_ex0);
    AXSQ_CHECK_HIP(hipGetLastError());
    k_refit_span<true><<<(unsigned)nspans, kRefitThreads, 0, s>>>(
        b->leaf_boxes, inner, lch, rch, par, rng, ex0, n_ex0, nchunks, ex1, n_ex1, cnt, ib);
    AXSQ_CHECK_HIP(hipGetLastError());
    if (nspans > 1) {  // one span holds the whole tree: no exits left
      k_refit_span<false><<<1, kRefitThreads, 0, s>>>(b->leaf_boxes, inner, lch, rch, par, rng,
                                                      ex1, n_ex1, 0, nullptr, nullptr, cnt, ib);
      AXSQ_CHECK_HIP(hipGetLastError());
    }
    k_emit<<<grid_for(inner), kBlock, 0, s>>>(b->leaf_boxes, ib, inner, lch, rch, rng, b->nodes,
                                              n32, b->ranges32, b->status);
    AXSQ_CHECK_HIP(hipGetLastError());
    return AXSQ_OK;
  };

  if (!b->host_params) {
    void* hp = nullptr;
    AXSQ_CHECK_HIP(hipHostMalloc(&hp, sizeof(BuildParams), hipHostMallocDefault));
    b->host_params = hp;
    AXSQ_CHECK_HIP(hipHostGetDevicePointer(&b->host_params_dev, hp, 0));
  }
  BuildParams* host_out = static_cast<BuildParams*>(b->host_params_dev);
  ProfRange pr("bvh_build", s);
  const int bits = dims == 2 ? 32 : 30;
  uint32_t *keys = ws.k0.as<uint32_t>(), *vals = ws.v0.as<uint32_t>();
  if (small) {
    // one tile-major histogram per sort pass: the first from the codes, the
    // others cleared by the transform and filled by the scatters
    const int32_t hs = small_sort_stride(n);
    const int npass = (bits + 7) / 8;
    const int64_t nhist = (int64_t)npass * 256 * hs;
    AXSQ_TRY(ws.hist.ensure(sizeof(uint32_t) * (size_t)nhist));
    uint32_t* hists = ws.hist.as<uint32_t>();
    const unsigned nkb = (unsigned)((n + kSmallSortTile - 1) / kSmallSortTile);
    auto enqueue = [&](hipStream_t q) -> int {
      if (f32)
        k_transform<true><<<(unsigned)nparts, kBlock, 0, q>>>(boxes_f32, dims, n_in, n, scale,
                                                              scaled, partials, hists + 256 * hs,
                                                              nhist - 256 * hs);
      else
        k_transform<false><<<(unsigned)nparts, kBlock, 0, q>>>(boxes, 3, n_in, n, scale, scaled,
                                                               partials, hists + 256 * hs,
                                                               nhist - 256 * hs);
      AXSQ_CHECK_HIP(hipGetLastError());
      k_mcodes_small<<<nkb, 1024, 0, q>>>(partials, nparts, scaled, n, dims, f32 ? 1 : 0,
                                          host_out, keys, vals, hists);
      AXSQ_CHECK_HIP(hipGetLastError());
      AXSQ_TRY(radix_sort_small_build(keys, vals, ws.k1.as<uint32_t>(), ws.v1.as<uint32_t>(), n,
                                      bits, hists, scaled, b->leaf_boxes, b->mcodes, b->leaf_ids,
                                      q));
      k_seg_local<<<(unsigned)nseg, 2 * kSegLeaves, 0, q>>>(
          b->mcodes, inner, lch, rch, par, rng, b->leaf_boxes, n, P, levels, seg, b->nodes, n32,
          b->ranges32, b->status, deferred, ndef);
      if (nseg > 1) {  // one workgroup of k_seg_local holds every node
        k_emit_seg<<<2048, kBlock, 0, q>>>(b->leaf_boxes, seg, P, rng, deferred, ndef, nseg,
                                           b->nodes, n32, b->ranges32);
        AXSQ_CHECK_HIP(hipGetLastError());
      }
      AXSQ_CHECK_HIP(hipGetLastError());
      return AXSQ_OK;
    };
    // the launches as one HIP graph, captured on a stream of our own and
    // replayed while the inputs, sizes and every buffer are the ones it holds
    // (AXSQ_BUILD_GRAPH=0: plain launches)
    const std::vector<uint64_t> key = {
        (uint64_t)(uintptr_t)(f32 ? (const void*)boxes_f32 : (const void*)boxes),
        (uint64_t)n_in, (uint64_t)dims, f32 ? 1u : 0u, (uint64_t)__builtin_bit_cast(uint64_t, scale),
        (uint64_t)(uintptr_t)n32, (uint64_t)(uintptr_t)b->nodes, (uint64_t)(uintptr_t)b->ranges32,
        (uint64_t)(uintptr_t)b->leaf_boxes, (uint64_t)(uintptr_t)b->leaf_ids,
        (uint64_t)(uintptr_t)b->mcodes, (uint64_t)(uintptr_t)b->status,
        (uint64_t)(uintptr_t)base, (uint64_t)(uintptr_t)hists, (uint64_t)(uintptr_t)keys,
        (uint64_t)(uintptr_t)vals, (uint64_t)(uintptr_t)ws.k1.as<uint32_t>(),
        (uint64_t)(uintptr_t)ws.v1.as<uint32_t>(), (uint64_t)(uintptr_t)host_out};
    const char* genv = getenv("AXSQ_BUILD_GRAPH");
    const bool use_graph = !(genv && genv[0] == '0');
    if (use_graph && !(b->build_graph && b->build_graph_key == key)) {
      if (b->build_graph) hipGraphExecDestroy(b->build_graph);
      b->build_graph = nullptr;
      b->build_graph_key.clear();
      hipStream_t cs = capture_stream();
      hipGraph_t graph = nullptr;
      if (cs && hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        const int rc = enqueue(cs);
        const hipError_t ec = hipStreamEndCapture(cs, &graph);
        if (rc == AXSQ_OK && ec == hipSuccess && graph) {
          hipGraphExec_t ex = nullptr;
          if (hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0) == hipSuccess) {
            b->build_graph = ex;
            b->build_graph_key = key;
          }
        }
        if (graph) hipGraphDestroy(graph);
        (void)hipGetLastError();  // a failed capture falls back to plain launches
      }
    }
    if (use_graph && b->build_graph) {
      AXSQ_CHECK_HIP(hipGraphLaunch(b->build_graph, s));
    } else {
      AXSQ_TRY(enqueue(s));
    }
  } else {
    if (f32)
      k_transform<true><<<(unsigned)nparts, kBlock, 0, s>>>(boxes_f32, dims, n_in, n, scale,
                                                            scaled, partials, nullptr, 0);
    else
      k_transform<false><<<(unsigned)nparts, kBlock, 0, s>>>(boxes, 3, n_in, n, scale, scaled,
                                                             partials, nullptr, 0);
    AXSQ_CHECK_HIP(hipGetLastError());
    k_bounds<<<1, 1024, 0, s>>>(partials, nparts, scaled, n, prm, host_out, f32 ? 1 : 0);
    AXSQ_CHECK_HIP(hipGetLastError());
    k_mcodes<<<grid_for(n), kBlock, 0, s>>>(scaled, n, prm, dims, f32 ? 1 : 0, keys, vals);
    AXSQ_CHECK_HIP(hipGetLastError());
    uint32_t *sk, *sv;
    AXSQ_TRY(radix_sort_pairs(keys, vals, ws.k1.as<uint32_t>(), ws.v1.as<uint32_t>(), n, bits,
                              ws, s, &sk, &sv));
    k_reorder<<<grid_for(n), kBlock, 0, s>>>(scaled, sv, sk, n, b->leaf_boxes, b->leaf_ids,
                                             b->mcodes);
    AXSQ_CHECK_HIP(hipGetLastError());
    k_karras<<<grid_for(inner), kBlock, 0, s>>>(b->mcodes, inner, lch, rch, par, rng, cnt, inner,
                                                n_ex1, 1);
    AXSQ_CHECK_HIP(hipGetLastError());
    AXSQ_TRY(climb_and_emit());
  }
  pr.stop();
  const BuildParams& hp = *static_cast<const BuildParams*>(b->host_params);
  AXSQ_CHECK_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < 6; ++k) b->bounds[k] = hp.bounds[k];
  if (small && hp.quirky) {
    // a leaf box with a signed-zero pair: unions may depend on the grouping,
    // so the boxes come from the climb, in the SEQ order
    AXSQ_TRY(fill_i32(cnt, 0, inner, s));
    AXSQ_TRY(fill_i32(n_ex1, 0, 1, s));
    AXSQ_TRY(climb_and_emit());
    AXSQ_CHECK_HIP(hipStreamSynchronize(s));
  }
  prof_flush();
  return AXSQ_OK;
}

}  // namespace axsq
