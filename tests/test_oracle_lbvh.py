"""The LBVH restatement (oracle/lbvh_oracle.hpp) against brute force and the
invariants the reference checks in LBVH::bvhSelfCheck (lbvh.cu:479-541)."""
import numpy as np
import pytest

from oracle.pyoracle import OracleLbvh
from tests.lbvh_cases import brute_pairs, segment_boxes, recipe_points


def self_check(t, boxes):
    n = len(t["ids"])
    codes, links, lp = t["codes"], t["links"], t["leaf_parent"]
    assert np.all(codes[1:] >= codes[:-1]), "Morton order"
    assert sorted(t["ids"].tolist()) == list(range(n))
    # children point back to their parent
    for i in range(1, n - 1):
        par = links[i, 0]
        p, right = par & 0x7FFFFFFF, par >> 31
        assert links[p, 2 if right else 1] == i
    for i in range(n):
        p, right = lp[i] & 0x7FFFFFFF, lp[i] >> 31
        assert links[p, 2 if right else 1] == (i | 0x80000000)
    # every node reachable, DFS stack bound, index ranges and boxes merge
    stack, seen, best = [0], 0, 1
    while stack:
        c = stack.pop()
        seen += 1
        if c >> 31:
            continue
        stack += [int(links[c, 1]), int(links[c, 2])]
        best = max(best, len(stack))
    assert seen == 2 * n - 1
    assert t["max_stack"] == best

    def rng(c):
        if c >> 31:
            leaf = c & 0x7FFFFFFF
            return leaf, leaf, boxes[t["ids"][leaf]]
        lo1, hi1, b1 = rng(int(links[c, 1]))
        lo2, hi2, b2 = rng(int(links[c, 2]))
        assert (min(lo1, lo2), max(hi1, hi2)) == (min(c, links[c, 3]), max(c, links[c, 3]))
        np.testing.assert_array_equal(t["bounds"][c, :6], b1)
        np.testing.assert_array_equal(t["bounds"][c, 6:], b2)
        return min(lo1, lo2), max(hi1, hi2), np.concatenate([np.minimum(b1[:3], b2[:3]), np.maximum(b1[3:], b2[3:])])

    import sys
    sys.setrecursionlimit(10000)
    rng(0)
    radix_prefix_check(t)
    return True


def radix_prefix_check(t):
    """Karras (2012) radix-tree rule, from its definition: an internal node
    covering sorted positions [lo, hi] splits where the highest bit in which
    the 64-bit keys (code << 32 | position) of lo and hi differ turns from 0
    to 1.  Its left child is node `split` (a leaf if split == lo) and its right
    child node `split + 1` (a leaf if split + 1 == hi)
    (KE/KittenGpuLBVH/lbvh.cu:68-150).  The split is found here by scanning
    the bit over the range, independently of the binary search the oracle
    uses."""
    n = len(t["codes"])
    if n < 2:
        return
    keys = (t["codes"].astype(np.uint64) << np.uint64(32)) | np.arange(n, dtype=np.uint64)
    links = t["links"]
    leaf = 0x80000000
    for c in range(n - 1):
        fence = int(links[c, 3])
        lo, hi = min(c, fence), max(c, fence)
        diff = int(keys[lo] ^ keys[hi])
        assert diff != 0
        bit = np.uint64(1) << np.uint64(diff.bit_length() - 1)
        ones = (keys[lo:hi + 1] & bit) != 0
        split = lo + int(np.argmax(ones)) - 1  # last position with the bit clear
        assert not ones[:split - lo + 1].any() and ones[split - lo + 1:].all()
        # every key in the range shares the bits above `bit`
        assert np.all((keys[lo:hi + 1] ^ keys[lo]) < (bit << np.uint64(1)))
        want_left = (split | leaf) if split == lo else split
        want_right = ((split + 1) | leaf) if split + 1 == hi else split + 1
        assert int(links[c, 1]) == want_left, (c, lo, hi, split)
        assert int(links[c, 2]) == want_right, (c, lo, hi, split)


def check_query(t, pairs, boxes):
    pos = np.empty(len(t["ids"]), np.int64)
    pos[t["ids"]] = np.arange(len(t["ids"]))
    # found object first, at the larger sorted position
    assert np.all(pos[pairs[:, 0]] > pos[pairs[:, 1]])
    got = np.sort(pairs.astype(np.int64), axis=1)
    key = got[:, 0] * len(boxes) + got[:, 1]
    assert len(np.unique(key)) == len(key), "duplicate pairs"
    want = brute_pairs(boxes)
    wkey = np.sort(want[:, 0] * len(boxes) + want[:, 1])
    assert np.array_equal(np.sort(key), wkey)


def test_reference_recipe_small(oracle_lib):
    boxes = recipe_points(3000, pad=0.01)
    o = OracleLbvh(boxes)
    t = o.get()
    assert self_check(t, boxes)
    check_query(t, o.query(), boxes)


def test_segments_with_empty_tips(oracle_lib):
    rng = np.random.default_rng(4)
    ns, nv = 40, 20
    roots = rng.uniform(0, 0.01, size=(ns, 3))
    pos = (roots[:, None, :] + np.linspace(0, 0.02, nv)[None, :, None] * np.array([0, -1.0, 0])
           + rng.normal(scale=2e-4, size=(ns, nv, 3))).reshape(-1, 3).astype(np.float32)
    off = np.arange(0, ns * nv + 1, nv)
    boxes = segment_boxes(pos, off, 1e-4)
    o = OracleLbvh(boxes)
    t = o.get()
    # empty boxes have NaN centres -> Morton code 0
    assert np.all(t["codes"][:ns] == 0)
    assert self_check(t, boxes)
    check_query(t, o.query(), boxes)
    # refit with moved boxes keeps the structure and matches brute force
    boxes2 = segment_boxes(pos + rng.normal(scale=5e-4, size=pos.shape).astype(np.float32), off, 1e-4)
    o.refit(boxes2)
    t2 = o.get()
    assert np.array_equal(t2["links"], t["links"])
    assert self_check(t2, boxes2)
    check_query(t2, o.query(), boxes2)


def test_duplicate_codes(oracle_lib):
    boxes = np.zeros((257, 6), np.float32)
    boxes[:, 3:] = 1.0  # all identical boxes -> identical codes, ties by index
    boxes[0] = [0, 0, 0, 2, 2, 2]
    o = OracleLbvh(boxes)
    t = o.get()
    assert self_check(t, boxes)
    check_query(t, o.query(cap=257 * 256), boxes)
