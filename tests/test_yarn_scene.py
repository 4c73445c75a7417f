"""YarnBall JSON scenes (the YBJsonPath setting) on the CPU: the host part
of kamisimulator_amd/host/yarn_scene.cpp through `kscr_solver_cli ybscene`.

The reference builds these scenes with YarnBall::buildFromJSON and
Sim::glueEndpoints (YarnBall/io/jsonBuilder.cpp:7-188).  Pins:
- the Sim::Sim defaults (YBSimulator.cu:81-112) and buildFromJSON's own
  (density 1, kStretch 5e5, kBend 0.1, transform diag(0.01)) for an empty
  scene, and every "simulation" key landing in its own parameter;
- the transform rows against the curve reader run with the same matrix
  (`curves2hair`, whose transform is checked in test_curve_io.py);
- closed curves glued end to end, and not when broken up;
- glueEndpoints on hand-built scenes whose outcome follows from the rule
  (closest vertex with |i - j| > 2, the lowest index on ties, both vertices
  unconnected, both moved to the midpoint), and the grid search against the
  reference's brute-force loop on a random scene;
- fixBorders and fixVertex on scenes whose pinned sets are worked out by
  hand."""
import json
import struct

import numpy as np
import pytest

from kamisimulator_amd import hairio

F32 = np.float32
IDENTITY = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]]


def _bcc(path, curves, kind=b"PL", closed=()):
    header = b"BCC" + bytes([0x44]) + kind + bytes([3, 1])
    header += struct.pack("<QQ", len(curves), sum(len(c) for c in curves)) + bytes(40)
    with open(path, "wb") as f:
        f.write(header)
        for k, c in enumerate(curves):
            f.write(struct.pack("<i", -len(c) if k in closed else len(c)))
            f.write(np.ascontiguousarray(c, F32).tobytes())


def _scene(tmp_path, curves, closed=(), **keys):
    _bcc(tmp_path / "c.bcc", curves, closed=closed)
    scene = {"curveFile": str(tmp_path / "c.bcc")}
    scene.update(keys)
    (tmp_path / "scene.json").write_text(json.dumps(scene))
    out = json.loads(hairio.cli("ybscene", tmp_path / "scene.json", tmp_path / "state.bin").stdout)
    raw = np.fromfile(tmp_path / "state.bin", np.uint8)
    n = int(raw[:4].view(np.int32)[0])
    f = raw[4:].view(F32)
    out["positions"] = f[:3 * n].reshape(n, 3)
    out["glued"] = f[3 * n:].reshape(n, 3)
    return out


def _line(n, start, step):
    return (np.asarray(start, np.float64) + np.arange(n)[:, None] * np.asarray(step, np.float64)).astype(F32)


def test_empty_scene_defaults(tmp_path):
    curves = [_line(5, (0, 0, 0), (1, 0, 0)), _line(4, (0, 1, 0), (0, 0, 1))]
    s = _scene(tmp_path, curves)
    p = s["params"]
    # Sim::Sim (YBSimulator.cu:81-112)
    assert p["gravity"] == pytest.approx([0, -9.8, 0]) and p["num_itr"] == 8
    for k, v in dict(max_h=4e-4, damping=1e-6, velocity_decay=0.2, radius=1e-4, barrier_thickness=8e-4,
                     acceleration_ratio=1, k_collision=1e-5, k_friction=5, friction_coeff=0.1, detection_scaler=1.2,
                     bvh_rebuild_period=0.1, external_force_multiplier=1, wind_peak=0, wind_sharpness=1).items():
        assert p[k] == pytest.approx(v, rel=1e-7), k
    assert (p["detection_period"], p["use_step_size_limit"], p["use_velocity_radius"],
            p["enable_hair_hair_coll"]) == (1, 0, 1, 0)
    # buildFromJSON's own defaults
    assert (s["density"], s["kStretch"], s["kBend"]) == (1.0, 5e5, pytest.approx(0.1, rel=1e-7))
    assert s["resample"] == 0 and s["breakUpClosedCurves"] == 0 and s["glueRadius"] == -1
    assert s["transform"] == pytest.approx([0.01, 0, 0, 0, 0, 0.01, 0, 0, 0, 0, 0.01, 0, 0, 0, 0, 1], rel=1e-7)
    want = np.concatenate(curves) * F32(0.01)
    np.testing.assert_array_equal(s["positions"], want)
    np.testing.assert_array_equal(s["glued"], want)
    assert s["strandOffsets"] == [0, 5, 9] and s["closed"] == [0, 0]
    assert s["connections"] == [-1] * 9 and s["pinned"] == []


def test_every_simulation_key(tmp_path):
    sim = {"density": 1300, "frictionCoeff": 0.3, "kStretch": 123.5, "kBend": 0.25, "kCollision": 2e-5,
           "kFriction": 7, "damping": 3e-6, "gravity": [1, 2, 3], "drag": 0.7, "maxTimeStep": 2e-4,
           "numIterations": 12, "detectionPeriod": 3, "detectionScaler": 1.5, "stepLimit": True,
           "velStepLimit": False, "bvhRebuildPeriod": 0.25}
    s = _scene(tmp_path, [_line(5, (0, 0, 0), (1, 0, 0))], simulation=sim, curveRadius=2e-3)
    p = s["params"]
    assert (s["density"], s["kStretch"], s["kBend"]) == (1300, 123.5, 0.25)
    want = dict(friction_coeff=0.3, k_collision=2e-5, k_friction=7, damping=3e-6, velocity_decay=0.7, max_h=2e-4,
                detection_scaler=1.5, bvh_rebuild_period=0.25)
    for k, v in want.items():
        assert p[k] == pytest.approx(v, rel=1e-7), k
    assert p["gravity"] == [1, 2, 3] and p["num_itr"] == 12 and p["detection_period"] == 3
    assert p["use_step_size_limit"] == 1 and p["use_velocity_radius"] == 0
    # curveRadius r: radius 0.05 r, barrier 2 (1 - 0.05) r, in float
    r = F32(2e-3)
    assert F32(p["radius"]) == F32(0.05) * r and F32(p["barrier_thickness"]) == F32(2 * (1 - F32(0.05))) * r
    # what the scene does not name keeps the defaults
    assert p["enable_hair_hair_coll"] == 0 and p["external_force_multiplier"] == 1


def test_transform_rows_match_the_curve_reader(tmp_path):
    rng = np.random.default_rng(5)
    curves = [rng.normal(size=(n, 3)).astype(F32) for n in (6, 4, 9)]
    rows = [[0.5, 0.1, 0.0, 1.0], [0.0, 2.0, 0.3, -1.0], [0.2, 0.0, 1.5, 0.25]]
    s = _scene(tmp_path, curves, transform=rows)
    M = np.eye(4, dtype=F32)
    M[:3, :] = np.asarray(rows, F32)
    assert s["transform"] == pytest.approx(M.T.ravel().tolist(), rel=1e-7)  # column-major
    args = ["curves2hair", tmp_path / "c.bcc", tmp_path / "c.bin", "0.001", 0, 0,
            ",".join(repr(float(v)) for v in M.T.ravel())]
    hairio.cli(*args)
    counts, cur, _ = hairio.read_hair(tmp_path / "c.bin")
    ref = np.concatenate([cur[k, :counts[k]] for k in range(len(curves))])
    np.testing.assert_array_equal(s["positions"], ref)
    # and the affine map itself, within float rounding
    P = np.concatenate(curves).astype(np.float64)
    np.testing.assert_allclose(s["positions"], P @ M[:3, :3].T.astype(np.float64) + M[:3, 3], rtol=0, atol=1e-6)


def test_closed_curves_are_glued_end_to_end(tmp_path):
    curves = [_line(5, (0, 0, 0), (1, 0, 0)), _line(6, (0, 2, 0), (1, 0.5, 0)), _line(4, (0, 5, 0), (0, 0, 1))]
    s = _scene(tmp_path, curves, closed=(1,))
    c = s["connections"]
    assert s["closed"] == [0, 1, 0] and c[5] == 10 and c[10] == 5
    assert sum(x >= 0 for x in c) == 2
    s = _scene(tmp_path, curves, closed=(1,), breakUpClosedCurves=True)
    assert s["closed"] == [0, 0, 0] and s["connections"] == [-1] * 15


MM = 2.0 ** -10  # about a millimetre, exact in float


def test_glue_endpoints_hand_cases(tmp_path):
    A = _line(5, (0, 0, 0), (MM, 0, 0))                          # vertices 0-4
    B = _line(5, (2 * MM, MM / 4, 0), (0, MM, 0))                # 5-9: B0 a quarter above A2
    D = _line(4, (2 * MM, -MM / 8, 0), (0, -MM, 0))              # 10-13: D0 an eighth below A2
    C = _line(4, (20 * MM, 0, 0), (0, 0, MM))                    # 14-17: far away
    # E: 18-23, its tip comes back next to its own vertex 20 (|i - j| = 3) and,
    # closer still, next to vertex 21 (|i - j| = 2, excluded)
    E = np.array([[0, 0, 10], [1, 0, 10], [2, 0, 10], [3, 0, 10], [3, 1, 10], [2.6, 0.3, 10]], F32) * F32(MM)
    radius = 0.75 * MM
    s = _scene(tmp_path, [A, B, D, C, E], transform=IDENTITY, glueEndpoints=radius)
    c = s["connections"]
    P0, G = s["positions"], s["glued"]
    # B0 (5) glues to A2 (2); both move to the midpoint
    assert c[2] == 5 and c[5] == 2
    mid = F32(0.5) * (P0[2] + P0[5])
    np.testing.assert_array_equal(G[2], mid)
    np.testing.assert_array_equal(G[5], mid)
    # D0 (10): its closest vertices are A2 and B0, both moved to the same
    # point and already glued: D0 stays unglued
    assert c[10] == -1
    # E's tip (23): vertex 21 is closer but |23 - 21| = 2; vertex 20 is taken
    d20 = np.linalg.norm(P0[23].astype(np.float64) - P0[20])
    d21 = np.linalg.norm(P0[23].astype(np.float64) - P0[21])
    assert d21 < d20 <= radius
    assert c[23] == 20 and c[20] == 23
    # nothing else is glued and nothing else moves
    assert sum(x >= 0 for x in c) == 4
    moved = np.flatnonzero(np.any(G != P0, axis=1))
    assert moved.tolist() == [2, 5, 20, 23]
    # a radius below every gap glues nothing; a negative one skips gluing
    for r in (MM / 16, -1.0):
        s = _scene(tmp_path, [A, B, D, C, E], transform=IDENTITY, glueEndpoints=r)
        assert s["connections"] == [-1] * 24 and np.array_equal(s["glued"], s["positions"])


def test_glue_endpoints_tie_takes_the_lowest_index(tmp_path):
    # the tip of T (vertex 14) lies exactly halfway between vertex 1 of L and vertex 1 of R
    L = _line(4, (-MM / 4, -MM, 0), (0, MM, 0))       # 0-3, vertex 1 at (-MM/4, 0, 0)
    R = _line(4, (MM / 4, -MM, 0), (0, MM, 0))        # 4-7, vertex 5 at (+MM/4, 0, 0)
    F = _line(4, (5 * MM, 0, 0), (MM, 0, 0))          # 8-11
    T = _line(3, (0, 0, 6 * MM), (0, 0, -3 * MM))     # 12-14, tip at the origin
    s = _scene(tmp_path, [L, R, F, T], transform=IDENTITY, glueEndpoints=MM / 2)
    c = s["connections"]
    # endpoints 0 and 3 come first: L0 and R0, and L3 and R3, are 0.5 MM apart (|i - j| > 2)
    assert c[0] == 4 and c[4] == 0 and c[3] == 7 and c[7] == 3
    assert c[14] == 1 and c[1] == 14
    assert sum(x >= 0 for x in c) == 6


def _brute_force_glue(P, off, conn, radius):
    """The loop of Sim::glueEndpoints (jsonBuilder.cpp:163-188), one thread,
    in float32."""
    P = P.copy()
    n = len(P)
    has_next = np.ones(n, bool)
    has_next[np.asarray(off[1:]) - 1] = False
    ends = [i for i in range(n) if (i > 0 and has_next[i - 1]) != has_next[i]]
    for i in ends:
        d = P - P[i]
        dist = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
        dist[max(0, i - 2):i + 3] = np.inf
        j = int(np.argmin(dist))  # first index of the minimum
        if dist[j] <= F32(radius) and conn[j] < 0 and conn[i] < 0:
            conn[j], conn[i] = i, j
            P[i] = P[j] = F32(0.5) * (P[i] + P[j])
    return P, conn


def test_glue_grid_search_equals_the_brute_force_loop(tmp_path):
    rng = np.random.default_rng(11)
    curves = []
    for _ in range(60):
        n = int(rng.integers(3, 12))
        start = rng.uniform(0, 40 * MM, 3)
        curves.append((start + np.cumsum(rng.normal(scale=MM, size=(n, 3)), 0)).astype(F32))
    radius = 2 * MM
    s = _scene(tmp_path, curves, transform=IDENTITY, glueEndpoints=radius)
    off = s["strandOffsets"]
    P, conn = _brute_force_glue(s["positions"], off, np.full(off[-1], -1), radius)
    print(f"{int((conn >= 0).sum()) // 2} glued pairs")
    assert (conn >= 0).sum() >= 10
    assert s["connections"] == conn.tolist()
    np.testing.assert_array_equal(s["glued"], P)


def test_fix_borders(tmp_path):
    strand = _line(11, (0, 0, 0), (MM, 0, 0))
    other = _line(4, (5 * MM, 2 * MM, 0), (0, MM, 0))   # y from 2 to 5 MM
    # max.x -= 1.5 MM, min.x += 2.5 MM, max.y -= 1.5 MM; pad 0.01 * radius (radius 1e-4)
    s = _scene(tmp_path, [strand, other], transform=IDENTITY,
               fixBorders=[1.5 * MM, 2.5 * MM, 1.5 * MM, 0, 0, 0])
    # x in [2.5 MM - 1e-6, 8.5 MM + 1e-6]: strand vertices 0-2 and 9-10 leave it;
    # y <= 5 MM + 1e-6 - 1.5 MM: the other strand's vertices at 4 and 5 MM leave it
    assert s["pinned"] == [0, 1, 2, 9, 10, 13, 14]
    # not six numbers: ignored
    s = _scene(tmp_path, [strand, other], transform=IDENTITY, fixBorders=[MM, MM, MM, MM, MM])
    assert s["pinned"] == []
    # the box is closed (Bound::contains): with no padding (curveRadius 0) and
    # no borders, vertices on its faces stay free
    s = _scene(tmp_path, [strand, other], transform=IDENTITY, curveRadius=0, fixBorders=[0] * 6)
    assert s["pinned"] == []


def test_fix_vertex(tmp_path):
    strand = _line(11, (0, 0, 0), (MM, 0, 0))
    s = _scene(tmp_path, [strand], transform=IDENTITY,
               fixVertex=[7, [3.4 * MM, 0.1 * MM, 0], [9 * MM, 0, 0, 1.5 * MM], [1, 2]])
    # index 7; closest to 3.4 MM: vertex 3; within 1.5 MM of 9 MM (strictly): 8, 9, 10; a short array: skipped
    assert s["pinned"] == [3, 7, 8, 9, 10]


def test_scene_errors(tmp_path):
    _bcc(tmp_path / "c.bcc", [_line(5, (0, 0, 0), (1, 0, 0))])
    cases = [({"curveFile": str(tmp_path / "missing.bcc")}, "cannot open"),
             ({"curveFile": str(tmp_path / "c.bcc"), "fixVertex": [5]}, "fixVertex index 5"),
             ({"curveFile": str(tmp_path / "c.bcc"), "simulation": {"density": "heavy"}}, "density"),
             ({"curveFile": str(tmp_path / "c.bcc"), "transform": {"a": 1}}, "not an array")]
    for scene, msg in cases:
        (tmp_path / "bad.json").write_text(json.dumps(scene))
        r = hairio.cli("ybscene", tmp_path / "bad.json", tmp_path / "state.bin", check=False)
        assert r.returncode == 1 and msg in r.stderr, (scene, r.stderr)


def test_curve_file_next_to_the_scene(tmp_path):
    (tmp_path / "sub").mkdir()
    _bcc(tmp_path / "sub" / "yarn.bcc", [_line(5, (0, 0, 0), (1, 0, 0))])
    (tmp_path / "sub" / "scene.json").write_text(json.dumps({"curveFile": "yarn.bcc"}))
    r = hairio.cli("ybscene", tmp_path / "sub" / "scene.json", tmp_path / "state.bin")
    assert json.loads(r.stdout)["strandOffsets"] == [0, 5]


def test_settings_name_the_scene(tmp_path):
    (tmp_path / "s.json").write_text(json.dumps({"YBJsonPath": "/some/scene.json"}))
    assert json.loads(hairio.cli("parse", tmp_path / "s.json").stdout)["YBJsonPath"] == "/some/scene.json"
