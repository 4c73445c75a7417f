"""Box sets for the LBVH tests.

recipe_points follows the reference's own recipe (Kitten::testLBVH,
KE/KittenGpuLBVH/lbvh.cu:554-565): srand(1), points (rand()/RAND_MAX)^3 in the
unit cube, each box padded by R = 0.001.  glibc's rand() is called through
ctypes, so the points are the reference's points on this platform."""
import numpy as np

from kamisimulator_amd import scenes


def recipe_points(n=100_000, pad=0.001):
    return scenes.lbvh_recipe_boxes(n, pad)


def segment_boxes(positions, offsets, pad):
    """Padded segment boxes of strands; tips get empty boxes (+inf, -inf)."""
    n = positions.shape[0]
    boxes = np.empty((n, 6), np.float32)
    boxes[:, :3] = np.inf
    boxes[:, 3:] = -np.inf
    has_next = np.ones(n, bool)
    has_next[np.asarray(offsets[1:]) - 1] = False
    i = np.nonzero(has_next)[0]
    p0, p1 = positions[i], positions[i + 1]
    boxes[i, :3] = np.minimum(p0, p1) - np.float32(pad)
    boxes[i, 3:] = np.maximum(p0, p1) + np.float32(pad)
    return boxes


def brute_pairs(boxes):
    """All strictly overlapping pairs (i < j), Bound::intersects semantics."""
    b = boxes.astype(np.float32)
    n = b.shape[0]
    out = []
    for i in range(n):
        ov = ~(np.any(b[i, 3:] <= b[i + 1:, :3], axis=1) | np.any(b[i, :3] >= b[i + 1:, 3:], axis=1))
        for j in np.nonzero(ov)[0]:
            out.append((i, i + 1 + j))
    return np.array(out, np.int64).reshape(-1, 2)


def query_stack_need(tree, box, t):
    """Deepest stack the device query of sorted object t (box = its box) uses
    on `tree` (an OracleLbvh.get() dict): the same visit rule, pushing the
    right child only when both child boxes overlap and neither is a leaf."""
    leaf = 0x80000000
    links, bounds = tree["links"], tree["bounds"]

    def hit(b):  # the device test: open intervals overlap
        return bool(np.all(b[3:] > box[:3]) and np.all(b[:3] < box[3:]))

    stack, deepest, node = [], 0, 0
    while node is not None:
        nxt = None
        lc, rc, fence = (int(v) for v in links[node, 1:4])
        if max(node, fence) > t:
            h0 = hit(bounds[node, :6]) and not (lc & leaf)
            h1 = hit(bounds[node, 6:]) and not (rc & leaf)
            if h0 and h1:
                stack.append(rc)
                deepest = max(deepest, len(stack))
                nxt = lc
            elif h0:
                nxt = lc
            elif h1:
                nxt = rc
        if nxt is None and stack:
            nxt = stack.pop()
        node = nxt
    return deepest
