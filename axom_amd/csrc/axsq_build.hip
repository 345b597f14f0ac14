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

#include "axsq_internal.hpp"

namespace axsq {

namespace {

struct BuildParams {
  double bounds[6];
  double minc[3];
  double inv[3];
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
  }
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
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; ++k) partials[6 * (int64_t)blockIdx.x + k] = s_red[k][0];
}

// Union of the valid scaled boxes; with none valid, the SEQ reduce
// (addBox from an invalid box, build_radix_tree.hpp:134-142) leaves the last
// box. Then inv_extent as in get_mcodes (:154-163), in float for a float
// tree (f32: extent and 1 / extent rounded to float, the zero test's
// threshold 1e-8 as a float).
__global__ __launch_bounds__(1024) void k_bounds(const double* __restrict__ partials,
                                                  int64_t nparts, const double* __restrict__ boxes,
                                                  int64_t n, BuildParams* __restrict__ prm,
                                                  int f32) {
  __shared__ double s_red[6][1024];
  double acc[6];
  invalid_box(acc);
  for (int64_t p = threadIdx.x; p < nparts; p += 1024) {
    for (int k = 0; k < 3; ++k) {
      const double a = partials[6 * p + k];
      if (a < acc[k]) acc[k] = a;
      const double b = partials[6 * p + 3 + k];
      if (b > acc[3 + k]) acc[3 + k] = b;
    }
  }
  for (int k = 0; k < 6; ++k) s_red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
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
    double b[6];
    for (int k = 0; k < 6; ++k) b[k] = s_red[k][0];
    if (!box_valid(b)) load_box(boxes + 6 * (n - 1), b);
    for (int k = 0; k < 6; ++k) prm->bounds[k] = b[k];
    for (int d = 0; d < 3; ++d) {
      prm->minc[d] = b[d];
      if (f32) {
        const float ext = (float)b[3 + d] - (float)b[d];
        prm->inv[d] = (fabsf(ext - 0.0f) <= (float)1.0e-8) ? 0.0 : (double)(1.0f / ext);
      } else {
        const double ext = b[3 + d] - b[d];
        prm->inv[d] = near_eq(ext, 0.0, 1.0e-8) ? 0.0 : 1.0 / ext;
      }
    }
  }
}

// hist0 (if given): the sort's first-pass tile histogram (low 8 bits).
// f32: a float tree, the normalised centroid computed in float.
__global__ __launch_bounds__(kBlock) void k_mcodes(const double* __restrict__ boxes, int64_t n,
                                                    const BuildParams* __restrict__ prm, int dims,
                                                    int f32, uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals,
                                                    uint32_t* __restrict__ hist0) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double b[6];
  load_box(boxes + 6 * i, b);
  double c[3], minc[3], inv[3];
  for (int d = 0; d < 3; ++d) {
    c[d] = 0.5 * (b[d] + b[3 + d]);
    minc[d] = prm->minc[d];
    inv[d] = prm->inv[d];
  }
  if (f32) {  // (c - min) * inv in float, then codes of that point over [0, 1]
    mcode_coords_f32(b, minc, inv, c);
    for (int d = 0; d < 3; ++d) {
      minc[d] = 0.0;
      inv[d] = 1.0;
    }
  }
  const uint32_t code = (dims == 2) ? morton2_of(c, minc, inv) : morton_of(c, minc, inv);
  keys[i] = code;
  vals[i] = (uint32_t)i;
  if (hist0) atomicAdd(&hist0[(i / kSmallSortTile) * 256 + (code & 255u)], 1u);
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

// build_tree, build_radix_tree.hpp:290-384
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
      box_add_box(acc, ob);
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
      box_add_box(acc, ob);
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
// FP32 nodes whose min corners round down and max corners round up.
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
  Node64 nd;
  int32_t lc = lchild[i], rc = rchild[i];
  if (lc >= inner) {
    load_box(leaf_boxes + 6 * (int64_t)(lc - inner), nd.l);
    lc = -(lc - inner + 1);
  } else {
    load_box(inner_boxes + 6 * (int64_t)lc, nd.l);
  }
  if (rc >= inner) {
    load_box(leaf_boxes + 6 * (int64_t)(rc - inner), nd.r);
    rc = -(rc - inner + 1);
  } else {
    load_box(inner_boxes + 6 * (int64_t)rc, nd.r);
  }
  nd.lc = lc;
  nd.rc = rc;
  nd.flags = (box_valid(nd.l) ? 1 : 0) | (box_valid(nd.r) ? 2 : 0);
  nd.lo = range[3 * i];
  nd.hi = range[3 * i + 1];
  nd.mid = range[3 * i + 2];
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
    ranges32[i] = make_int4(nd.lo, nd.hi, nd.mid, 0);
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
  // flag buffer holds: scaled boxes | partials | params | lchild | rchild |
  // parents | counters | leaf ranges | inner boxes
  const size_t off_part = sizeof(double) * 6 * n;
  const size_t off_prm = off_part + sizeof(double) * 6 * nparts;
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
  const size_t total = off_ex + sizeof(int32_t) * (nchunks + 1 + nchunks * kBlock + n);
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

  ProfRange pr("bvh_build", s);
  // inputs of up to kSmallSortMax boxes: the first sort pass's histogram is
  // built with the codes, the others by the scatters
  const int bits = dims == 2 ? 32 : 30;
  const bool small_sort = n <= kSmallSortMax;
  const int64_t sort_tiles = (n + kSmallSortTile - 1) / kSmallSortTile;
  const int64_t nhist = small_sort ? 4 * 256 * sort_tiles : 0;
  uint32_t* hists = nullptr;
  if (small_sort) {
    AXSQ_TRY(ws.hist.ensure(sizeof(uint32_t) * (size_t)nhist));
    hists = ws.hist.as<uint32_t>();
  }
  if (f32)
    k_transform<true><<<(unsigned)nparts, kBlock, 0, s>>>(boxes_f32, dims, n_in, n, scale, scaled,
                                                          partials, hists, nhist);
  else
    k_transform<false><<<(unsigned)nparts, kBlock, 0, s>>>(boxes, 3, n_in, n, scale, scaled,
                                                           partials, hists, nhist);
  AXSQ_CHECK_HIP(hipGetLastError());
  k_bounds<<<1, 1024, 0, s>>>(partials, nparts, scaled, n, prm, f32 ? 1 : 0);
  AXSQ_CHECK_HIP(hipGetLastError());
  uint32_t *keys = ws.k0.as<uint32_t>(), *vals = ws.v0.as<uint32_t>();
  k_mcodes<<<grid_for(n), kBlock, 0, s>>>(scaled, n, prm, dims, f32 ? 1 : 0, keys, vals, hists);
  AXSQ_CHECK_HIP(hipGetLastError());
  uint32_t *sk, *sv;
  if (small_sort)
    AXSQ_TRY(radix_sort_small_prehist(keys, vals, ws.k1.as<uint32_t>(), ws.v1.as<uint32_t>(), n,
                                      bits, hists, s, &sk, &sv));
  else
    AXSQ_TRY(radix_sort_pairs(keys, vals, ws.k1.as<uint32_t>(), ws.v1.as<uint32_t>(), n, bits,
                              ws, s, &sk, &sv));
  k_reorder<<<grid_for(n), kBlock, 0, s>>>(scaled, sv, sk, n, b->leaf_boxes, b->leaf_ids,
                                           b->mcodes);
  AXSQ_CHECK_HIP(hipGetLastError());
  k_karras<<<grid_for(inner), kBlock, 0, s>>>(b->mcodes, inner, lch, rch, par, rng, cnt, inner,
                                              n_ex1, 1);
  AXSQ_CHECK_HIP(hipGetLastError());
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
  AXSQ_CHECK_HIP(hipGetLastError());
  k_emit<<<grid_for(inner), kBlock, 0, s>>>(b->leaf_boxes, ib, inner, lch, rch, rng, b->nodes,
                                            fp_mode == AXSQ_FP32_TRAVERSAL ? b->nodes32 : nullptr,
                                            b->ranges32,
                                            b->status);
  AXSQ_CHECK_HIP(hipGetLastError());
  pr.stop();
  AXSQ_CHECK_HIP(hipMemcpyAsync(b->bounds, prm->bounds, sizeof(double) * 6,
                                hipMemcpyDeviceToHost, s));
  AXSQ_CHECK_HIP(hipStreamSynchronize(s));
  prof_flush();
  return AXSQ_OK;
}

}  // namespace axsq
