"""Properties the BVH build relies on, checked on the CPU.

1. Node boxes as unions over leaf ranges. The SEQ refit gives a node
   V(right).addBox(V(left)) (build_radix_tree.hpp:536-571; BoundingBox::addBox,
   BoundingBox.hpp:486-508). As an operation on runs of leaves,
   comb(L, R) = R.addBox(L) must be associative bit for bit, so that any split
   of a range into consecutive pieces gives the SEQ tree's box. The one
   exception is a box with a coordinate whose min and max are zeros of
   opposite sign, which the build detects and sends to the exact climb.
2. The 8-way Karras searches (karras_node_wide) find the range end and split
   of build_radix_tree.hpp:290-384's doubling, binary and ceil(float(l)/div)
   searches.
3. The pieces k_emit_seg takes for a leaf range (left pieces below level 9,
   a run of level-9 nodes, right pieces) cover the range exactly, in order.
"""
import itertools
import math
import random
import struct

import numpy as np

BIG = 1.7e308


def _bits(b):
    return tuple(struct.pack("<d", x) for x in b)


def _valid(b, nd):
    return all(not (b[d] > b[nd + d]) for d in range(nd))


def _add_box(a, o, nd):  # a.addBox(o)
    a = list(a)
    if _valid(a, nd):
        if _valid(o, nd):
            for p in (o[:nd], o[nd:]):
                for d in range(nd):
                    if p[d] < a[d]:
                        a[d] = p[d]
                    if p[d] > a[nd + d]:
                        a[nd + d] = p[d]
    else:
        a = list(o)
    return tuple(a)


def _comb(left, right, nd):
    return _add_box(right, left, nd)


def _signed_zero_pair(b, nd):
    return any(b[d] == b[nd + d] and _bits([b[d]]) != _bits([b[nd + d]]) for d in range(nd))


def test_box_union_associative_except_signed_zero_pairs_1d():
    vals = [-1.0, -0.0, 0.0, 1.0, 2.0, BIG, -BIG]
    boxes = [(x, y) for x in vals for y in vals]  # inverted (invalid) ones included
    broken = 0
    for a, b, c in itertools.product(boxes, repeat=3):
        x = _comb(_comb(a, b, 1), c, 1)
        y = _comb(a, _comb(b, c, 1), 1)
        if _bits(x) != _bits(y):
            broken += 1
            assert any(_signed_zero_pair(t, 1) for t in (a, b, c)), (a, b, c)
    assert broken > 0  # the exception is real: the build must detect it


def test_box_union_associative_except_signed_zero_pairs_2d():
    rng = random.Random(5)
    vals = [-1.0, -0.0, 0.0, 1.0, 2.0, BIG, -BIG]
    for _ in range(60_000):
        a, b, c = (tuple(rng.choice(vals) for _ in range(4)) for _ in range(3))
        if any(_signed_zero_pair(t, 2) for t in (a, b, c)):
            continue
        assert _bits(_comb(_comb(a, b, 2), c, 2)) == _bits(_comb(a, _comb(b, c, 2), 2))


def test_scaled_box_signed_zero_pair_only_from_negative_zero_input():
    """box_scale (BoundingBox.hpp:547-561): mid +/- (max - min) * sf / 2. A
    coordinate [lo, hi] of a valid box ends with equal zeros of opposite sign
    only when lo = hi = -0.0."""
    for lo, hi in itertools.product([-0.0, 0.0, -1e-300, 1e-300, 1.0], repeat=2):
        if lo > hi:
            continue
        for sf in (1.0, 1.000123, 2.0):
            mid = 0.5 * (lo + hi)
            r = (hi - lo) * (sf * 0.5)
            a, b = mid - r, mid + r
            pair = a == b and _bits([a]) != _bits([b])
            assert pair == (_bits([lo]) == _bits([-0.0]) and _bits([hi]) == _bits([-0.0]))


def _delta_fn(mc, inner):
    def delta(a, b):
        if b < 0 or b > inner:
            return -1
        ex = int(mc[a]) ^ int(mc[b])
        tie = ex == 0
        if tie:
            ex = a ^ b
        c = 32 - ex.bit_length() if ex else 32
        return c + 32 if tie else c
    return delta


def _karras_reference(mc, inner, i):
    delta = _delta_fn(mc, inner)
    d = -1 if delta(i, i + 1) - delta(i, i - 1) < 0 else 1
    dmin = delta(i, i - d)
    lmax = 2
    while delta(i, i + lmax * d) > dmin:
        lmax *= 2
    l, s_ = 0, lmax // 2
    while s_ >= 1:
        if delta(i, i + (l + s_) * d) > dmin:
            l += s_
        s_ //= 2
    j = i + l * d
    dnode = delta(i, j)
    s, div = 0, 2.0
    step = math.ceil(float(np.float32(l)) / div)
    while True:
        if delta(i, i + (s + step) * d) > dnode:
            s += step
        if step == 1:
            break
        div *= 2
        step = math.ceil(float(np.float32(l)) / div)
    return min(i, j), max(i, j), i + s * d + (d if d < 0 else 0)


def _karras_wide(mc, inner, i):
    delta = _delta_fn(mc, inner)

    def last_pass(lo, hi, d, thr):
        while hi > lo:
            span = hi - lo
            p = [lo + (span * (k + 1) + 7) // 8 for k in range(7)]
            nlo, nhi = lo, hi
            for x in p:
                if delta(i, i + x * d) > thr:
                    nlo = max(nlo, x)
                else:
                    nhi = min(nhi, x - 1)
            lo, hi = nlo, nhi
        return lo

    dp, dm = delta(i, i + 1), delta(i, i - 1)
    d = -1 if dp - dm < 0 else 1
    dmin = dm if d > 0 else dp
    lmax = 1 << 19
    for k in range(18, 0, -1):
        if not delta(i, i + (1 << k) * d) > dmin:
            lmax = 1 << k
    l = last_pass(lmax // 2 if lmax >= 4 else 0, lmax - 1, d, dmin)
    j = i + l * d
    s = last_pass(0, l - 1, d, delta(i, j))
    return min(i, j), max(i, j), i + s * d + (d if d < 0 else 0)


def test_wide_karras_search_equals_reference_loops():
    rng = np.random.default_rng(3)
    for trial in range(24):
        n = int(rng.integers(2, 1500))
        mc = np.sort(rng.integers(0, [8, 64, 1 << 10, 1 << 30][trial % 4], n)).astype(np.uint32)
        for i in range(n - 1):
            assert _karras_reference(mc, n - 1, i) == _karras_wide(mc, n - 1, i), (trial, i)


def _deferred_pieces(a, b, P, log=9):
    l, r, nr = a + P, b + 1 + P, 0
    for _ in range(log):
        if not l < r:
            break
        if l & 1:
            l += 1
        if r & 1:
            nr += 1
            r -= 1
        l >>= 1
        r >>= 1
    run = list(range(l, r)) if l < r else []
    left, right = [], [None] * nr
    l, r, cr = a + P, b + 1 + P, 0
    for _ in range(log):
        if not l < r:
            break
        if l & 1:
            left.append(l)
            l += 1
        if r & 1:
            r -= 1
            right[nr - 1 - cr] = r
            cr += 1
        l >>= 1
        r >>= 1
    assert len(left) <= 9 and nr <= 9 and len(run) <= 8 * 64
    return left + run + right


def test_deferred_node_pieces_cover_the_range_in_order():
    rng = random.Random(11)
    for _ in range(4000):
        P = 1 << rng.randint(10, 18)
        n = rng.randint(P // 2 + 1, P)
        a = rng.randint(0, n - 2)
        b = rng.randint(a + 1, n - 1)
        leaves = []
        for x in _deferred_pieces(a, b, P):
            lo, lev = x, 0
            while lo < P:
                lo <<= 1
                lev += 1
            leaves += range(lo - P, lo - P + (1 << lev))
        assert leaves == list(range(a, b + 1))
