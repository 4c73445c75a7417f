"""Device LBVH (kscr_lbvh_*) against the restatement: Morton codes, sorted
order, tree links, node boxes, leaf parents, root box and stack bound are
compared bit for bit; the query against the restatement's pair set, on the
reference's own testLBVH recipe (100k points, srand(1), R = 0.001)."""
import time

import numpy as np
import pytest
import torch  # before the library loads the HIP runtime (device buffers of the in-place test)

from kamisimulator_amd._lib import KscrError, Lbvh
from oracle.pyoracle import OracleLbvh
from tests.lbvh_cases import query_stack_need, segment_boxes, recipe_points

pytestmark = pytest.mark.gpu


def _compare(boxes, label, cap=None, slots=None):
    print(f"[gpu-lbvh] {label}: n={len(boxes)}", flush=True)
    g = Lbvh(device=0)
    if slots is not None:
        g.set_fit_slots(slots)
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


def _refit_compare(g, o, boxes):
    g.refit(boxes)
    o.refit(boxes)
    tg, to = g.get(), o.get()
    assert np.array_equal(tg["bounds"].view(np.uint32), to["bounds"].view(np.uint32))
    assert np.array_equal(tg["root"].view(np.uint32), to["root"].view(np.uint32))
    assert tg["max_stack"] == to["max_stack"]


@pytest.mark.parametrize("n", [2, 3, 511, 512, 513, 2049, 131072, 131073])
def test_sizes_around_the_small_tree_bound(n):
    """Trees over at most 131072 boxes sort with one launch per digit pass and
    fit by rounds (the nodes inside each 512-leaf tile, then the nodes that
    cross a tile boundary, from LDS); larger ones take the multi-launch sort
    and the height-ordered refit.  Both against the restatement at the tile
    (512), radix-tile (2048) and small-tree bounds, with empty boxes (+inf,
    -inf) among the objects, then a refit with moved boxes."""
    rng = np.random.default_rng(n)
    P = rng.random((n, 3)).astype(np.float32)
    boxes = np.concatenate([P - 1e-3, P + 1e-3], 1).astype(np.float32)
    if n > 64:
        boxes[5::61] = [np.inf, np.inf, np.inf, -np.inf, -np.inf, -np.inf]
    g, o = _compare(boxes, f"uniform n={n}")
    moved = boxes + np.repeat(rng.normal(scale=2e-3, size=(n, 3)), 2, axis=0).reshape(n, 6).astype(np.float32)
    _refit_compare(g, o, moved)


@pytest.mark.parametrize("n", [5000, 140_000])
def test_nan_coordinates(n):
    """Boxes with a NaN coordinate (a NaN centre gives Morton cell 0; unions
    pass NaN on through the same selects as glm's min/max) on a small tree
    and a large one: the tree, node boxes, root box and stack bound equal the
    restatement after the build and after a refit.  (Such a box overlaps
    everything in the reference's intersects, so the query is not part of
    this check.)"""
    rng = np.random.default_rng(n + 1)
    P = rng.random((n, 3)).astype(np.float32)
    boxes = np.concatenate([P - 1e-3, P + 1e-3], 1).astype(np.float32)
    boxes[11::89, 1] = np.nan
    boxes[13::97, 5] = np.nan
    g = Lbvh(device=0)
    g.compute(boxes)
    o = OracleLbvh(boxes)
    tg, to = g.get(), o.get()
    for k in ("codes", "ids", "links", "leaf_parent"):
        assert np.array_equal(tg[k], to[k]), k
    for k in ("bounds", "root"):
        assert np.array_equal(tg[k].view(np.uint32), to[k].view(np.uint32)), k
    assert tg["max_stack"] == to["max_stack"]
    assert np.isnan(tg["bounds"]).any()  # NaN reached the node boxes
    moved = boxes + np.float32(1e-3)
    _refit_compare(g, o, moved)


def test_crossing_nodes_fit_from_global_memory():
    """With fewer LDS slots than tile-crossing nodes (about 1.8k on the
    recipe; 0 slots here) the small-tree fit runs its rounds from global
    memory: the same tree, boxes and stack bound, also after a refit."""
    boxes = recipe_points(100_000)
    g, o = _compare(boxes, "testLBVH recipe, crossing nodes from global memory", slots=0)
    rng = np.random.default_rng(2)
    moved = boxes + np.repeat(rng.normal(scale=1e-3, size=(len(boxes), 3)), 2, axis=0).reshape(-1, 6).astype(
        np.float32)
    _refit_compare(g, o, moved)
    g.set_fit_slots(4096)  # back to LDS for the same tree
    _refit_compare(g, o, boxes)
    for bad in (-1, 4097):
        with pytest.raises(KscrError, match="INVALID_ARGUMENT"):
            g.set_fit_slots(bad)


def test_device_entry_points_read_boxes_in_place():
    """kscr_lbvh_compute_device and kscr_lbvh_refit_device read the caller's
    boxes where they lie, as the reference's LBVH keeps the pointer compute
    was given: the tree built from one device buffer, refit from a second
    (moved boxes) and queried, equals the restatement at each step; the
    refit and the query follow the second buffer, and a host-array refit
    afterwards takes its own upload."""
    torch.cuda.set_device(0)
    n = 30_000
    boxes = recipe_points(n)
    rng = np.random.default_rng(12)
    moved = (boxes + np.repeat(rng.normal(scale=2e-3, size=(n, 3)), 2, axis=0).reshape(n, 6)).astype(np.float32)
    d_a = torch.from_numpy(boxes).to("cuda:0")
    d_b = torch.from_numpy(moved).to("cuda:0")
    cap = 64 * n
    d_pairs = torch.empty((cap, 2), dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    g = Lbvh(device=0)
    o = OracleLbvh(boxes)
    g.compute_device(d_a.data_ptr(), n)
    torch.cuda.synchronize()
    tg, to = g.get(), o.get()
    for k in ("codes", "ids", "links", "leaf_parent"):
        assert np.array_equal(tg[k], to[k]), k
    assert np.array_equal(tg["bounds"].view(np.uint32), to["bounds"].view(np.uint32))
    g.refit_device(d_b.data_ptr())
    cnt = g.query_device(d_pairs.data_ptr(), cap)
    o.refit(moved)
    tg, to = g.get(), o.get()
    assert np.array_equal(tg["bounds"].view(np.uint32), to["bounds"].view(np.uint32))
    op = o.query(cap)
    gp = d_pairs[:cnt].cpu().numpy()
    key = lambda p: np.sort(p[:, 0].astype(np.int64) * n + p[:, 1])
    assert cnt == len(op) and np.array_equal(key(gp), key(op))
    print(f"[gpu-lbvh] in-place device boxes: {cnt} pairs after the refit from the second buffer", flush=True)
    _refit_compare(g, o, boxes)  # host arrays: uploaded into the handle's own buffer


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
