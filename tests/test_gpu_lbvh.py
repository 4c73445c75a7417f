"""Device LBVH (kscr_lbvh_*) against the restatement: Morton codes, sorted
order, tree links, node boxes, leaf parents, root box and stack bound are
compared bit for bit; the query against the restatement's pair set, on the
reference's own testLBVH recipe (100k points, srand(1), R = 0.001)."""
import time

import numpy as np
import pytest

from kamisimulator_amd._lib import Lbvh
from oracle.pyoracle import OracleLbvh
from tests.lbvh_cases import query_stack_need, segment_boxes, recipe_points

pytestmark = pytest.mark.gpu


def _compare(boxes, label, cap=None):
    print(f"[gpu-lbvh] {label}: n={len(boxes)}", flush=True)
    g = Lbvh(device=0)
    t0 = time.perf_counter()
    g.compute(boxes)
    t1 = time.perf_counter()
    o = OracleLbvh(boxes)
    tg, to = g.get(), o.get()
    for k in ("codes", "ids", "links", "leaf_parent"):
        assert np.array_equal(tg[k], to[k]), k
    for k in ("bounds", "root"):
        assert np.array_equal(tg[k].view(np.uint32), to[k].view(np.uint32)), k
    assert tg["max_stack"] == to["max_stack"]
    cap = 200 * len(boxes) if cap is None else cap
    gp, cnt = g.query(cap)
    op = o.query(cap)
    assert cnt == len(op) and cnt <= cap
    key = lambda p: np.sort(p[:, 0].astype(np.int64) * len(boxes) + p[:, 1])
    assert np.array_equal(key(gp), key(op))  # same oriented pairs
    print(f"[gpu-lbvh] {label}: {cnt} pairs, build+download {1e3 * (t1 - t0):.1f} ms, stack {tg['max_stack']}",
          flush=True)
    return g, o


def test_reference_recipe_100k():
    _compare(recipe_points(100_000), "testLBVH recipe")


def test_segments_and_refit():
    rng = np.random.default_rng(9)
    ns, nv = 2000, 50
    roots = rng.uniform(0, 0.05, size=(ns, 3))
    pos = (roots[:, None, :] + np.linspace(0, 0.05, nv)[None, :, None] * np.array([0, -1.0, 0.2])
           + rng.normal(scale=3e-4, size=(ns, nv, 3))).reshape(-1, 3).astype(np.float32)
    off = np.arange(0, ns * nv + 1, nv)
    boxes = segment_boxes(pos, off, 1e-4)
    g, o = _compare(boxes, "segments")
    boxes2 = segment_boxes(pos + rng.normal(scale=4e-4, size=pos.shape).astype(np.float32), off, 1e-4)
    g.refit(boxes2)
    o.refit(boxes2)
    tg, to = g.get(), o.get()
    assert np.array_equal(tg["bounds"].view(np.uint32), to["bounds"].view(np.uint32))
    assert tg["max_stack"] == to["max_stack"]


def test_ties_and_one_object():
    boxes = np.zeros((5000, 6), np.float32)
    boxes[:, 3:] = 1.0
    boxes[::7] = [0, 0, 0, 2, 2, 2]
    g = Lbvh(device=0)
    g.compute(boxes)
    o = OracleLbvh(boxes)
    for k in ("codes", "ids", "links", "leaf_parent"):
        assert np.array_equal(g.get()[k], o.get()[k]), k
    one = np.array([[0, 0, 0, 1, 1, 1]], np.float32)
    g.compute(one)
    pairs, cnt = g.query(16)
    assert cnt == 0


def test_two_pass_query_above_2_20_objects():
    """Above 2^20 objects the query is two passes: a first pass walking two
    queries per lane with a short LDS stack, and a redo pass that finishes
    the queries needing more.  Boxes spanning the whole scene overlap every
    node on their way down, so their queries need a stack as deep as the tree
    and go to the redo pass; the pair set must still be the restatement's,
    with no pair lost or repeated."""
    rng = np.random.default_rng(13)
    n = (1 << 20) + 150_000
    pts = rng.uniform(0, 1, size=(n, 3)).astype(np.float32)
    small = np.concatenate([pts - 2e-3, pts + 2e-3], axis=1).astype(np.float32)
    big = np.array([[-0.5, -0.5, -0.5, 1.5, 1.5, 1.5],
                    [-1.0, -1.0, -1.0, 1.1, 1.1, 1.1],
                    [-0.1, -0.1, -0.1, 2.0, 2.0, 2.0]], np.float32)
    boxes = np.concatenate([small[:300_000], big[:1], small[300_000:700_000], big[1:], small[700_000:]])
    g, o = _compare(boxes, "two-pass query", cap=8_000_000)
    tree = o.get()
    pos = {int(obj): k for k, obj in enumerate(tree["ids"])}
    need = query_stack_need(tree, boxes[300_000], pos[300_000])
    print(f"[gpu-lbvh] stack needed by a whole-scene query: {need}", flush=True)
    assert need > 12  # deeper than the first pass's stack: the redo pass ran


def test_queries_deeper_than_the_first_pass_stack():
    """Boxes spanning the whole scene overlap every node on their way down,
    so their queries need a stack as deep as the tree; at 300k objects the
    query is one pass with the full stack.  The pair set must still be the
    restatement's, with no pair lost or repeated."""
    rng = np.random.default_rng(12)
    pts = rng.uniform(0, 1, size=(300_000, 3)).astype(np.float32)
    small = np.concatenate([pts - 2e-3, pts + 2e-3], axis=1).astype(np.float32)
    big = np.array([[-0.5, -0.5, -0.5, 1.5, 1.5, 1.5],      # centre at the middle
                    [-1.0, -1.0, -1.0, 1.1, 1.1, 1.1],      # centre near the low corner
                    [-0.1, -0.1, -0.1, 2.0, 2.0, 2.0]], np.float32)
    boxes = np.concatenate([small[:100_000], big[:1], small[100_000:200_000], big[1:], small[200_000:]])
    g, o = _compare(boxes, "whole-scene boxes")
    tree = o.get()
    pos = {int(obj): k for k, obj in enumerate(tree["ids"])}
    need = [query_stack_need(tree, boxes[b], pos[b]) for b in (100_000, 200_001, 200_002)]
    print(f"[gpu-lbvh] stack needed by the whole-scene queries: {need}", flush=True)
    assert max(need) > 16  # so the redo pass ran
