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
                                          double r[6], int32_t& lc, int32_t& rc, int32_t& flags) {
  const double2* p = reinterpret_cast<const double2*>(nodes + i);
  const double2 a0 = p[0], a1 = p[1], a2 = p[2], a3 = p[3], a4 = p[4], a5 = p[5];
  const int4 c = *reinterpret_cast<const int4*>(p + 6);
  l[0] = a0.x; l[1] = a0.y; l[2] = a1.x; l[3] = a1.y; l[4] = a2.x; l[5] = a2.y;
  r[0] = a3.x; r[1] = a3.y; r[2] = a4.x; r[3] = a4.y; r[4] = a5.x; r[5] = a5.y;
  lc = c.x;
  rc = c.y;
  flags = c.z;
}

__device__ __forceinline__ void load_node(const Node32* __restrict__ nodes, int32_t i, float l[6],
                                          float r[6], int32_t& lc, int32_t& rc, int32_t& flags) {
  const float4* p = reinterpret_cast<const float4*>(nodes + i);
  const float4 a0 = p[0], a1 = p[1], a2 = p[2], a3 = p[3];
  l[0] = a0.x; l[1] = a0.y; l[2] = a0.z; l[3] = a0.w; l[4] = a1.x; l[5] = a1.y;
  r[0] = a1.z; r[1] = a1.w; r[2] = a2.x; r[3] = a2.y; r[4] = a2.z; r[5] = a2.w;
  lc = __float_as_int(a3.x);
  rc = __float_as_int(a3.y);
  flags = __float_as_int(a3.z);
}

template <class T>
__device__ __forceinline__ bool valid6(const T b[6]) {
  return !(b[0] > b[3]) && !(b[1] > b[4]) && !(b[2] > b[5]);
}

// Per-lane traversal stack. The first kLdsStack entries live in LDS laid out
// [depth][lane] so that 64 lanes at 64 different depths still hit 64
// different banks; deeper entries (trees deeper than kLdsStack levels, rare)
// go to a private array. A private array indexed per lane would be placed in
// scratch, where lanes at different depths touch different cache lines.
// traverse_vote: a leaf-check step runs when at least this percentage of the
// walking lanes want one
#ifndef AXSQ_VOTE_LEAF_PCT
#define AXSQ_VOTE_LEAF_PCT 50
#endif
#ifndef AXSQ_LDS_STACK
#define AXSQ_LDS_STACK 32
#endif
constexpr int kLdsStack = AXSQ_LDS_STACK;
constexpr int kWaveLanes = 64;

// LDS-qualified word: accesses compile to ds_read/ds_write, not flat ops.
typedef __attribute__((address_space(3))) int32_t lds_i32;

// LDS: entries per lane kept in LDS (the rest of the kStack in the private
// array). A kernel whose walks stay shallow (point queries) keeps fewer, so
// that its blocks' LDS lets more waves in.
template <int LDS = kLdsStack>
struct LaneStackT {
  static_assert(LDS > 0 && LDS <= kStack, "LDS entries");
  lds_i32* lds;  // &tile[0][lane] of this wave's [LDS][64] tile
  int32_t spill[kStack - LDS > 0 ? kStack - LDS : 1];
  __device__ __forceinline__ int32_t get(int32_t sp) const {
    if (sp < LDS) return lds[sp * kWaveLanes];
    return spill[sp - LDS];
  }
  __device__ __forceinline__ void set(int32_t sp, int32_t v) {
    if (sp < LDS)
      lds[sp * kWaveLanes] = v;
    else
      spill[sp - LDS] = v;
  }
};
using LaneStack = LaneStackT<>;

// LDS words a block of `threads` lanes needs for its stacks.
template <int LDS = kLdsStack>
constexpr int stack_lds_words(int threads) { return threads * LDS; }

// This lane's stack view into a block-wide __shared__ array of
// stack_lds_words<LDS>(blockDim.x) words.
template <int LDS = kLdsStack>
__device__ __forceinline__ lds_i32* lane_stack_base(int32_t* block_stack) {
  const int wave = threadIdx.x / kWaveLanes;
  const int lane = threadIdx.x % kWaveLanes;
  return (lds_i32*)(block_stack + wave * (LDS * kWaveLanes) + lane);
}

// Resumable walk state: the stack lives in the LaneStack, the rest here.
// The walk can only stop at the top of its outer loop, where no leaf is held.
struct Walk {
  int32_t cur;     // next node (kBarrier when done)
  int32_t sp;      // top of stack
  int32_t visits;  // inner nodes visited so far
  int32_t steps;   // inner visits + leaf actions so far
};

template <class Stack>
__device__ __forceinline__ Walk walk_start(Stack& todo) {
  todo.set(0, kBarrier);
  return Walk{0, 0, 0, 0};
}

enum WalkStatus { kWalkDone = 0, kWalkSuspended = 1, kWalkError = -1 };

// pred(box) -> bool; pref(lbox, rbox) -> true to visit the right child first;
// leaf(pos) runs the leaf action. Runs until the walk ends (kWalkDone), or
// until `budget` steps have been taken at the top of the outer loop
// (kWalkSuspended, resumable with the same state), or kWalkError after a
// stack overflow. `ninner` bounds node links and visits: a well-formed tree
// visits each inner node at most once, so the guards only turn a malformed
// tree into an error instead of a stray access or an endless walk.
template <class NodeT, class Scalar, class Stack, class Pred, class Pref, class Leaf>
__device__ __forceinline__ int traverse(const NodeT* __restrict__ nodes, int32_t ninner,
                                        Stack& todo, Walk& w, int32_t budget, Pred&& pred,
                                        Pref&& pref, Leaf&& leaf) {
  int32_t sp = w.sp, cur = w.cur, visits = w.visits, steps = w.steps;
  int32_t found = 0;
  int status = kWalkDone;
  while (cur != kBarrier) {
    if (steps >= budget) {
      status = kWalkSuspended;
      break;
    }
    while (cur >= 0) {
      if (cur >= ninner || visits >= ninner) return kWalkError;
      Scalar lb[6], rb[6];
      int32_t lc, rc, flags;
      load_node(nodes, cur, lb, rb, lc, rc, flags);
      ++visits;
      ++steps;
      const bool in_l = (flags & 1) ? pred(lb) : false;
      const bool in_r = (flags & 2) ? pred(rb) : false;
      if (!in_l && !in_r) {
        cur = todo.get(sp);
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
          if (sp >= kStack) return kWalkError;
          todo.set(sp, rc);
        }
      }
      if (cur < 0 && !(found < 0)) {
        found = cur;
        if (cur != kBarrier) {
          cur = todo.get(sp);
          sp--;
        }
      }
    }
    while (found < 0 && found != kBarrier) {
      if (-found - 1 > ninner) return kWalkError;  // ninner + 1 leaves
      leaf(-found - 1);
      ++steps;
      found = cur;
      if (cur < 0 && cur != kBarrier) {
        if (sp < 0) return kWalkError;
        cur = todo.get(sp);
        sp--;
      }
    }
    found = 0;
  }
  w.sp = sp;
  w.cur = cur;
  w.visits = visits;
  w.steps = steps;
  return status;
}

// The same walk, scheduled by vote across the wave. Each lane runs exactly the
// sequence of node visits and leaf actions of traverse() above, but the wave
// no longer runs the inner (node) loop until every lane has found its next
// leaf and then the leaf loop until every lane is through its leaves: at each
// step the lanes whose next action is a leaf action and those whose next
// action is a node visit are counted, and the wave runs one leaf action (for
// the first kind) if they are at least half of the lanes still walking, else
// one node visit (for the second kind). A lane's own order of actions, and so
// every pruning decision, is unchanged; only which lanes wait changes. With
// walks of uneven shape in one wave this keeps about half the lanes busy in
// each step where the loop-by-loop schedule keeps a fifth busy. All lanes of
// the wave that are still running must call it together.
template <class NodeT, class Scalar, class Stack, class Pred, class Pref, class Leaf>
__device__ __forceinline__ int traverse_vote(const NodeT* __restrict__ nodes, int32_t ninner,
                                             Stack& todo, Walk& w, int32_t budget, Pred&& pred,
                                             Pref&& pref, Leaf&& leaf) {
  enum { kInner = 0, kLeaf = 1, kIdle = 2 };
  int32_t sp = w.sp, cur = w.cur, visits = w.visits, steps = w.steps;
  int32_t found = 0;
  int status = kWalkDone;
  int phase = kIdle;
  // top of the outer loop of traverse(): done, out of budget, or a node visit next
  auto top = [&]() {
    if (cur == kBarrier) {
      phase = kIdle;
    } else if (steps >= budget) {
      status = kWalkSuspended;
      phase = kIdle;
    } else if (cur < 0) {  // only reached from a malformed saved state
      status = kWalkError;
      phase = kIdle;
    } else {
      phase = kInner;
    }
  };
  // end of a leaf-loop check: leave the leaf loop when no leaf is held
  auto leaf_loop_test = [&]() {
    if (!(found < 0 && found != kBarrier)) {
      found = 0;
      top();
    }
  };
  top();
  for (;;) {
    const uint64_t want_node = __ballot(phase == kInner);
    const uint64_t want_leaf = __ballot(phase == kLeaf);
    if ((want_node | want_leaf) == 0) break;
    const int n_node = __popcll(want_node), n_leaf = __popcll(want_leaf);
    const bool leaf_step = n_leaf > 0 && 100 * n_leaf >= AXSQ_VOTE_LEAF_PCT * (n_node + n_leaf);
    if (leaf_step) {
      if (phase == kLeaf) {
        if (-found - 1 > ninner) {  // ninner + 1 leaves
          status = kWalkError;
          phase = kIdle;
        } else {
          leaf(-found - 1);
          ++steps;
          found = cur;
          if (cur < 0 && cur != kBarrier) {
            if (sp < 0) {
              status = kWalkError;
              phase = kIdle;
            } else {
              cur = todo.get(sp);
              sp--;
            }
          }
          if (phase == kLeaf) leaf_loop_test();
        }
      }
    } else if (phase == kInner) {
      if (cur >= ninner || visits >= ninner) {
        status = kWalkError;
        phase = kIdle;
      } else {
        Scalar lb[6], rb[6];
        int32_t lc, rc, flags;
        load_node(nodes, cur, lb, rb, lc, rc, flags);
        ++visits;
        ++steps;
        const bool in_l = (flags & 1) ? pred(lb) : false;
        const bool in_r = (flags & 2) ? pred(rb) : false;
        bool overflow = false;
        if (!in_l && !in_r) {
          cur = todo.get(sp);
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
            if (sp >= kStack)
              overflow = true;
            else
              todo.set(sp, rc);
          }
        }
        if (overflow) {
          status = kWalkError;
          phase = kIdle;
        } else {
          if (cur < 0 && !(found < 0)) {
            found = cur;
            if (cur != kBarrier) {
              cur = todo.get(sp);
              sp--;
            }
          }
          if (cur < 0) {  // the inner loop ends: on to the held leaves
            phase = kLeaf;
            leaf_loop_test();
          }
        }
      }
    }
  }
  w.sp = sp;
  w.cur = cur;
  w.visits = visits;
  w.steps = steps;
  return status;
}

// One-shot walk to completion; returns visited inner nodes or -1. lds_stack
// is this lane's base from lane_stack_base<LDS>.
template <class NodeT, class Scalar, int LDS = kLdsStack, class Pred, class Pref, class Leaf>
__device__ __forceinline__ int32_t traverse(const NodeT* __restrict__ nodes, int32_t ninner,
                                            lds_i32* lds_stack, Pred&& pred, Pref&& pref,
                                            Leaf&& leaf) {
  LaneStackT<LDS> todo;
  todo.lds = lds_stack;
  Walk w = walk_start(todo);
  const int st = traverse<NodeT, Scalar>(nodes, ninner, todo, w, 0x7fffffff, pred, pref, leaf);
  return st == kWalkDone ? w.visits : -1;
}

// The same one-shot walk on the vote schedule (every lane of the wave that is
// still running calls it together).
template <class NodeT, class Scalar, int LDS = kLdsStack, class Pred, class Pref, class Leaf>
__device__ __forceinline__ int32_t traverse_voted(const NodeT* __restrict__ nodes, int32_t ninner,
                                                  lds_i32* lds_stack, Pred&& pred, Pref&& pref,
                                                  Leaf&& leaf) {
  LaneStackT<LDS> todo;
  todo.lds = lds_stack;
  Walk w = walk_start(todo);
  const int st =
      traverse_vote<NodeT, Scalar>(nodes, ninner, todo, w, 0x7fffffff, pred, pref, leaf);
  return st == kWalkDone ? w.visits : -1;
}

}  // namespace axsq
