"""Closed, consistently oriented (outward) triangle meshes for the body SDF
tests, generated from seeds: a box, an icosphere and a torus, optionally
with jittered vertices and an offset."""
import numpy as np


def box(lo=(-0.03, -0.05, -0.02), hi=(0.04, 0.06, 0.025)):
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    v = np.array([[lo[0] if i & 1 == 0 else hi[0], lo[1] if i & 2 == 0 else hi[1], lo[2] if i & 4 == 0 else hi[2]]
                  for i in range(8)], np.float32)
    # two triangles per face, counter-clockwise seen from outside
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    t = []
    for a, b, c, d in quads:
        t += [(a, b, c), (a, c, d)]
    return v, np.array(t, np.int32)


def icosphere(level=2, radius=0.05, centre=(0.0, 0.0, 0.0)):
    p = (1 + 5 ** 0.5) / 2
    v = [(-1, p, 0), (1, p, 0), (-1, -p, 0), (1, -p, 0), (0, -1, p), (0, 1, p), (0, -1, -p), (0, 1, -p),
         (p, 0, -1), (p, 0, 1), (-p, 0, -1), (-p, 0, 1)]
    v = [np.array(x, np.float64) / np.linalg.norm(x) for x in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        mid = {}

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                x = v[a] + v[b]
                v.append(x / np.linalg.norm(x))
                mid[k] = len(v) - 1
            return mid[k]

        g = []
        for a, b, c in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            g += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = g
    v = np.array(v) * radius + np.asarray(centre)
    return v.astype(np.float32), np.array(f, np.int32)


def torus(n_major=24, n_minor=12, major=0.05, minor=0.018):
    v = []
    for i in range(n_major):
        u = 2 * np.pi * i / n_major
        for j in range(n_minor):
            w = 2 * np.pi * j / n_minor
            r = major + minor * np.cos(w)
            v.append((r * np.cos(u), r * np.sin(u), minor * np.sin(w)))
    t = []
    for i in range(n_major):
        for j in range(n_minor):
            a = i * n_minor + j
            b = ((i + 1) % n_major) * n_minor + j
            c = ((i + 1) % n_major) * n_minor + (j + 1) % n_minor
            d = i * n_minor + (j + 1) % n_minor
            t += [(a, b, c), (a, c, d)]
    return np.array(v, np.float32), np.array(t, np.int32)


def jitter(v, seed, scale=2e-4, offset=(0.0, 0.0, 0.0)):
    rng = np.random.default_rng(seed)
    return (v.astype(np.float64) + rng.uniform(-scale, scale, v.shape) + np.asarray(offset)).astype(np.float32)


def cases():
    """(name, vertices, triangles, dim) of the golden and parity cases."""
    bv, bt = box()
    iv, it = icosphere(2)
    tv, tt = torus()
    return [
        ("box", bv, bt, (10, 12, 8)),
        ("icosphere", jitter(iv, 1, offset=(0.01, -0.02, 0.005)), it, (12, 16, 10)),
        ("torus", jitter(tv, 2), tt, (14, 14, 8)),
    ]


def query_points(v, n, seed):
    """Points in the vertex box padded by 20 %, plus the vertices and edge
    midpoints' neighbourhood (features where ties and sign changes happen)."""
    rng = np.random.default_rng(seed)
    lo, hi = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
    pad = 0.2 * (hi - lo)
    p = rng.uniform(lo - pad, hi + pad, (n, 3))
    near = v[rng.integers(0, len(v), n // 4)].astype(np.float64) + rng.normal(0, 1e-3, (n // 4, 3))
    return np.concatenate([p, near])
