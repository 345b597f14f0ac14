// axsq_traverse.hpp -- per-lane stack traversal shared by the query kernels.
//
// Control flow follows the reference's "while-while" traversal
// (src/axom/spin/internal/linear_bvh/bvh_traverse.hpp:66-154) step for step:
// both children are tested when a node is visited, the nearer-first swap is
// applied only when both pass, the far child is pushed, and a found leaf is
// held back until a second leaf (or the end of the stack) is reached, then
// the held leaves are processed in order. Leaves and inner nodes taken from
// the stack are not tested again. Keeping this order makes every stateful
// query (the signed-distance pruning) see exactly the reference's sequence.
//
// Differences: node links index 128-byte records (the reference doubles the
// index into a two-box array), and a push past 64 entries stops the query
// and reports it instead of writing past the stack.
#pragma once

#include "axsq_device.hpp"

namespace axsq {

// Node access for the two node formats.
__device__ __forceinline__ void load_node(const Node64* __restrict__ nodes, int32_t i, double l[6],
                                          double r[6], int32_t& lc, int32_t& rc) {
  const double2* p = reinterpret_cast<const double2*>(nodes + i);
  const double2 a0 = p[0], a1 = p[1], a2 = p[2], a3 = p[3], a4 = p[4], a5 = p[5];
  const int2 c = *reinterpret_cast<const int2*>(p + 6);
  l[0] = a0.x; l[1] = a0.y; l[2] = a1.x; l[3] = a1.y; l[4] = a2.x; l[5] = a2.y;
  r[0] = a3.x; r[1] = a3.y; r[2] = a4.x; r[3] = a4.y; r[4] = a5.x; r[5] = a5.y;
  lc = c.x;
  rc = c.y;
}

__device__ __forceinline__ void load_node(const Node32* __restrict__ nodes, int32_t i, float l[6],
                                          float r[6], int32_t& lc, int32_t& rc) {
  const float4* p = reinterpret_cast<const float4*>(nodes + i);
  const float4 a0 = p[0], a1 = p[1], a2 = p[2], a3 = p[3];
  l[0] = a0.x; l[1] = a0.y; l[2] = a0.z; l[3] = a0.w; l[4] = a1.x; l[5] = a1.y;
  r[0] = a1.z; r[1] = a1.w; r[2] = a2.x; r[3] = a2.y; r[4] = a2.z; r[5] = a2.w;
  lc = __float_as_int(a3.x);
  rc = __float_as_int(a3.y);
}

template <class T>
__device__ __forceinline__ bool valid6(const T b[6]) {
  return !(b[0] > b[3]) && !(b[1] > b[4]) && !(b[2] > b[5]);
}

// pred(box) -> bool; pref(lbox, rbox) -> true to visit the right child first;
// leaf(pos) runs the leaf action. Returns the number of visited inner nodes,
// or -1 after a stack overflow. `ninner` bounds node links and visits: a
// well-formed tree visits each inner node at most once, so the guards only
// turn a malformed tree into an error instead of a stray access or an
// endless walk.
template <class NodeT, class Scalar, class Pred, class Pref, class Leaf>
__device__ __forceinline__ int32_t traverse(const NodeT* __restrict__ nodes, int32_t ninner,
                                            Pred&& pred, Pref&& pref, Leaf&& leaf) {
  int32_t todo[kStack];
  int32_t sp = 0;
  todo[0] = kBarrier;
  int32_t found = 0;
  int32_t cur = 0;
  int32_t visits = 0;
  while (cur != kBarrier) {
    while (cur >= 0) {
      if (cur >= ninner || visits >= ninner) return -1;
      Scalar lb[6], rb[6];
      int32_t lc, rc;
      load_node(nodes, cur, lb, rb, lc, rc);
      ++visits;
      const bool in_l = valid6(lb) ? pred(lb) : false;
      const bool in_r = valid6(rb) ? pred(rb) : false;
      if (!in_l && !in_r) {
        cur = todo[sp];
        sp--;
      } else {
        cur = in_l ? lc : rc;
        if (in_l && in_r) {
          if (pref(lb, rb)) {
            const int32_t t = cur;
            cur = rc;
            rc = t;
          }
          sp++;
          if (sp >= kStack) return -1;
          todo[sp] = rc;
        }
      }
      if (cur < 0 && !(found < 0)) {
        found = cur;
        if (cur != kBarrier) {
          cur = todo[sp];
          sp--;
        }
      }
    }
    while (found < 0 && found != kBarrier) {
      if (-found - 1 > ninner) return -1;  // ninner + 1 leaves
      leaf(-found - 1);
      found = cur;
      if (cur < 0 && cur != kBarrier) {
        if (sp < 0) return -1;
        cur = todo[sp];
        sp--;
      }
    }
    found = 0;
  }
  return visits;
}

}  // namespace axsq
