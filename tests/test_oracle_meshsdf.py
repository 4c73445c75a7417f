"""The body SDF oracle (oracle/meshsdf_oracle.hpp) pinned on the CPU.

1. Against the reference itself: tests/golden/mesh_sdf_vectors.npz holds what
   the reference's own Kami::SDF (src/utilities/Geometry.cpp with the vendored
   Discregrid TriangleMeshDistance and CubicLagrangeDiscreteGrid) and the node
   loop of PNMesh::StoreSDFQueries compute for three closed meshes
   (tests/golden/make_mesh_sdf_vectors.py).  The grid's node values, the
   sampled phi and grad, and the signed distances at query points must agree
   bit for bit.  The nearest triangle may differ at ties (shared vertices and
   edges): there the distance must still agree exactly.
2. Against mathematics: the shape functions form a partition of unity, are
   1 at their own node and 0 at the other 31, and their derivatives match
   finite differences; the grid interpolates any cubic polynomial exactly
   with its gradient (which pins the node positions, the cell-to-node map and
   the cell lookup); the signed distance to a box and to an icosphere matches
   the closed forms.
"""
import os

import numpy as np
import pytest

import mesh_cases
from oracle import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "mesh_sdf_vectors.npz")
NAMES = ["box", "icosphere", "torus"]


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


@pytest.mark.parametrize("name", NAMES)
def test_signed_distance_matches_reference(oracle_lib, golden, name):
    g = lambda k: golden[f"{name}_{k}"]  # noqa: E731
    sd, near, tri = po.mesh_distance(g("vertices"), g("triangles"), g("query"))
    np.testing.assert_array_equal(sd, g("sd"))
    np.testing.assert_allclose(near.ravel(), g("near"), rtol=0, atol=1e-15)
    # where the triangles differ the query is at a tie: same distance, and
    # the two nearest points coincide
    assert np.mean(tri == g("tri")) > 0.5


@pytest.mark.parametrize("name", NAMES)
def test_body_grid_matches_reference(oracle_lib, golden, name):
    g = lambda k: golden[f"{name}_{k}"]  # noqa: E731
    v = g("vertices")
    phi, grad, gmin, gmax, coef = po.mesh_sdf(v, g("triangles"), g("dim"), with_coef=True)
    np.testing.assert_array_equal(coef, g("coef"))
    np.testing.assert_array_equal(phi.ravel(), g("phi"))
    # the node at the grid centre (torus) has no radial direction: NaN in both
    np.testing.assert_array_equal(grad.ravel(), g("grad"))
    np.testing.assert_array_equal(gmin, v.min(0))
    np.testing.assert_array_equal(gmax, v.max(0))


def _abscissae():
    a = []
    for c in range(8):
        a.append([1 if c & 1 else -1, 1 if c & 2 else -1, 1 if c & 4 else -1])
    for axis in range(3):
        for e in range(4):
            if axis == 0:
                sb, sc = (1 if e & 2 else -1), (1 if e & 1 else -1)
            elif axis == 1:
                sb, sc = (1 if e & 1 else -1), (1 if e & 2 else -1)
            else:
                sb, sc = (1 if e & 2 else -1), (1 if e & 1 else -1)
            b, c = [x for x in range(3) if x != axis]
            for h in range(2):
                p = [0.0, 0.0, 0.0]
                p[axis] = -1 / 3 if h == 0 else 1 / 3
                p[b], p[c] = sb, sc
                a.append(p)
    return np.array(a, np.float64)


def test_shape_functions(oracle_lib):
    nodes = _abscissae()
    for j, xi in enumerate(nodes):
        N, _ = po.cubic_shape(xi)
        want = np.zeros(32)
        want[j] = 1
        np.testing.assert_allclose(N, want, atol=1e-15)
    rng = np.random.default_rng(0)
    h = 1e-6
    for xi in rng.uniform(-1, 1, (20, 3)):
        N, dN = po.cubic_shape(xi)
        assert abs(N.sum() - 1) < 1e-14
        np.testing.assert_allclose(dN.sum(0), 0, atol=1e-13)
        for a in range(3):
            e = np.zeros(3)
            e[a] = h
            fd = (po.cubic_shape(xi + e)[0] - po.cubic_shape(xi - e)[0]) / (2 * h)
            np.testing.assert_allclose(dN[:, a], fd, rtol=0, atol=1e-8)


def test_grid_reproduces_cubic_polynomials(oracle_lib):
    v, _ = mesh_cases.icosphere(1, radius=0.04, centre=(0.01, 0.0, -0.02))
    dim = (5, 7, 4)
    nodes = po.cubic_nodes(v, dim)
    lo = v.min(0).astype(np.float64) - 0.1
    hi = v.max(0).astype(np.float64) + 0.1
    assert nodes.min(0) == pytest.approx(lo) and nodes.max(0) == pytest.approx(hi)
    rng = np.random.default_rng(1)
    c = rng.normal(size=10)

    def f(p):
        x, y, z = p[:, 0], p[:, 1], p[:, 2]
        return (c[0] + c[1] * x + c[2] * y * z + c[3] * x * x * y + c[4] * z ** 3 + c[5] * x * y * z
                + c[6] * y ** 2 * x + c[7] * z * x * x + c[8] * y ** 3 + c[9] * x ** 3)

    def grad_f(p):
        x, y, z = p[:, 0], p[:, 1], p[:, 2]
        gx = c[1] + 2 * c[3] * x * y + c[5] * y * z + c[6] * y ** 2 + 2 * c[7] * z * x + 3 * c[9] * x ** 2
        gy = c[2] * z + c[3] * x * x + c[5] * x * z + 2 * c[6] * y * x + 3 * c[8] * y ** 2
        gz = c[2] * y + 3 * c[4] * z ** 2 + c[5] * x * y + c[7] * x * x
        return np.stack([gx, gy, gz], 1)

    pts = rng.uniform(lo, hi, (500, 3))
    val, grad = po.cubic_interpolate(v, dim, f(nodes), pts)
    np.testing.assert_allclose(val, f(pts), rtol=0, atol=1e-12)
    np.testing.assert_allclose(grad, grad_f(pts), rtol=0, atol=1e-10)


def test_box_closed_form(oracle_lib):
    lo, hi = np.array([-0.03, -0.05, -0.02]), np.array([0.04, 0.06, 0.025])
    v, t = mesh_cases.box(lo, hi)
    rng = np.random.default_rng(2)
    p = rng.uniform(lo - 0.05, hi + 0.05, (2000, 3))
    sd, near, _ = po.mesh_distance(v, t, p)
    lo32, hi32 = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
    q = np.maximum(lo32 - p, p - hi32)
    outside = np.linalg.norm(np.maximum(q, 0), axis=1)
    inside = np.minimum(q.max(1), 0)
    # the quadratic-form distance cancels: d2 carries ~eps |p - v0|^2
    np.testing.assert_allclose(sd, outside + inside, rtol=0, atol=1e-13)
    np.testing.assert_allclose(np.abs(sd), np.linalg.norm(p - near, axis=1), rtol=1e-9, atol=1e-15)


def test_icosphere_bounds(oracle_lib):
    r = 0.05
    v, t = mesh_cases.icosphere(3, radius=r)
    rng = np.random.default_rng(3)
    p = rng.uniform(-2 * r, 2 * r, (2000, 3))
    sd, _, _ = po.mesh_distance(v, t, p)
    exact = np.linalg.norm(p, axis=1) - r
    # the inscribed polyhedron lies within its sagitta of the sphere
    edge = np.linalg.norm(v[t[:, 0]].astype(np.float64) - v[t[:, 1]], axis=1).max()
    sag = r - np.sqrt(r * r - (edge / np.sqrt(3)) ** 2)
    assert np.all(sd >= exact - 1e-9) and np.all(sd <= exact + sag + 1e-9)
    np.testing.assert_array_equal(np.sign(sd[np.abs(exact) > sag]), np.sign(exact[np.abs(exact) > sag]))


def test_planted_mistake_fails_the_golden(oracle_lib, golden):
    """A plausible slip (the triangles' winding reversed, which flips every
    pseudo-normal) changes the signed distances: the golden comparison
    catches it."""
    g = lambda k: golden["icosphere_" + k]  # noqa: E731
    sd, _, _ = po.mesh_distance(g("vertices"), g("triangles")[:, ::-1], g("query"))  # reversed winding
    assert not np.array_equal(sd, g("sd"))
