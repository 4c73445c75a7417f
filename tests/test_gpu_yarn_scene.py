"""YarnBall JSON scenes (YBJsonPath) end to end on the GPU: the host solver's
Construct builds the simulation from the scene (kamisimulator_amd/host/
yarn_scene.cpp) and SolveNextStep advances it; the FP32 restatement set up
from the scene's host description (`kscr_solver_cli ybscene`: rest positions,
curves, parameters, connections, pins, glued positions) must give the same
words.  The reference: YarnBall::buildFromJSON (YarnBall/io/jsonBuilder.cpp)
as StableCosseratRodsSolver::Construct calls it
(StableCosseratRodsSolver.cpp:44-45)."""
import json
import struct

import numpy as np
import pytest

from kamisimulator_amd import hairio
from kamisimulator_amd._lib import default_params
from oracle.pyoracle import OracleSim
from tests import parity

pytestmark = pytest.mark.gpu
F32 = np.float32
DT = F32(1 / 600)


def _bcc(path, curves, closed=()):
    header = b"BCC" + bytes([0x44]) + b"PL" + bytes([3, 1])
    header += struct.pack("<QQ", len(curves), sum(len(c) for c in curves)) + bytes(40)
    with open(path, "wb") as f:
        f.write(header)
        for k, c in enumerate(curves):
            f.write(struct.pack("<i", -len(c) if k in closed else len(c)))
            f.write(np.ascontiguousarray(c, F32).tobytes())


def _read_state(path):
    raw = np.fromfile(path, np.uint8)
    n = int(raw[:4].view(np.int32)[0])
    f = raw[4:].view(F32)
    return f[:3 * n].reshape(n, 3), f[3 * n:].reshape(n, 3)


def _run_scene(tmp_path, scene, steps):
    (tmp_path / "scene.json").write_text(json.dumps(scene))
    desc = json.loads(hairio.cli("ybscene", tmp_path / "scene.json", tmp_path / "desc.bin").stdout)
    rest, glued = _read_state(tmp_path / "desc.bin")
    (tmp_path / "settings.json").write_text(json.dumps({"YBJsonPath": str(tmp_path / "scene.json"),
                                                         "EnableSDF": True, "NumItr": 3, "VelocityDecay": 9.0}))
    r = hairio.cli("run", tmp_path / "settings.json", "-", "-", steps, repr(float(DT)), tmp_path / "out.bin")
    refs = json.loads(r.stdout)
    assert refs["rootVertices"] == [] and refs["strandRefs"] == [] and refs["numVerts"] == len(rest)
    assert "YarnBall scene" in r.stderr
    got_x, got_v = _read_state(tmp_path / "out.bin")
    # the restatement, set up the way buildFromJSON sets up its Sim
    p = default_params(**desc["params"])
    o = OracleSim(p, rest, np.asarray(desc["strandOffsets"], np.int32), precision=32)
    o.configure(desc["density"])
    o.set_kstretch(desc["kStretch"])
    o.set_kbend(desc["kBend"])
    conn = np.asarray(desc["connections"], np.int32)
    if (conn >= 0).any():
        o.set_connections(conn)
    o.pin(desc["pinned"])
    o.upload()
    o.set_pos_vel(glued)
    o.set_threads(8)
    for _ in range(steps):
        o.advance(DT)
    op, ov = o.get_pos_vel()
    assert np.isfinite(got_x).all() and np.isfinite(got_v).all()
    parity.check("yarn scene", {"x": got_x, "v": got_v}, {"x": op, "v": ov}, {"x": glued})
    return desc, rest, glued, got_x


def test_bcc_scene_with_glue_rings_borders_and_pins(tmp_path):
    rows = []
    x = np.arange(41, dtype=np.float64)
    for r in range(6):
        rows.append(np.stack([x, r + 0.3 * np.sin(0.8 * x), 0.3 * np.cos(0.8 * x + r)], 1))
    verticals = [np.stack([np.full(7, x0 + 0.05), np.arange(7) - 0.4, np.zeros(7)], 1) for x0 in (5.0, 15.0, 25.0)]
    th = np.linspace(0, 2 * np.pi, 21)[:-1]
    ring = np.stack([50 + 2 * np.cos(th), 2 + 2 * np.sin(th), np.zeros(20)], 1)
    curves = [c.astype(F32) for c in rows + verticals + [ring]]
    _bcc(tmp_path / "knit.bcc", curves, closed=(len(curves) - 1,))
    scene = {"curveFile": "knit.bcc", "curveRadius": 2e-3, "glueEndpoints": 0.005,
             "fixBorders": [0.015, 0.015, 0, 0, 0, 0], "fixVertex": [[0.2, 0.0, 0.0], 3],
             "simulation": {"gravity": [0, -9.8, 0], "numIterations": 10, "drag": 0.5, "kStretch": 2e4,
                            "kBend": 0.05, "density": 0.5, "damping": 1e-5, "maxTimeStep": 5e-4}}
    print(f"[gpu-yarn] {len(curves)} curves, {sum(map(len, curves))} vertices, 60 steps", flush=True)
    desc, rest, glued, x = _run_scene(tmp_path, scene, 60)
    conn = np.asarray(desc["connections"])
    n_ring = len(rest) - 20
    assert conn[n_ring] == len(rest) - 1  # the ring is closed
    assert (conn >= 0).sum() >= 2 + 2 * 3  # and the verticals are glued to the rows
    assert len(desc["pinned"]) > 10 and 3 in desc["pinned"]
    # pinned vertices stay where the scene put them; free ones fall
    pinned = np.asarray(desc["pinned"])
    np.testing.assert_array_equal(x[pinned], glued[pinned])
    free = np.setdiff1d(np.arange(len(rest)), pinned)
    assert (x[free, 1] < glued[free, 1]).mean() > 0.5


def test_obj_scene_resampled_with_a_closed_loop(tmp_path):
    t = np.linspace(0, 4 * np.pi, 30)
    open_curve = np.stack([t, np.sin(t), 0.2 * np.cos(2 * t)], 1)
    th = np.linspace(0, 2 * np.pi, 17)[:-1]
    loop = np.stack([3 * np.cos(th), -4 + 3 * np.sin(th), 0.5 * np.sin(2 * th)], 1)
    lines = [f"v {a:.6f} {b:.6f} {c:.6f}" for a, b, c in np.concatenate([open_curve, loop])]
    lines.append("l " + " ".join(str(i + 1) for i in range(30)))
    lines.append("l " + " ".join(str(31 + i) for i in range(16)) + " 31")
    (tmp_path / "yarn.obj").write_text("\n".join(lines) + "\n")
    scene = {"curveFile": str(tmp_path / "yarn.obj"), "resampleLength": 0.004, "fixVertex": [0, [0.0, -0.01, 0.0]],
             "simulation": {"kStretch": 5e3, "kBend": 0.02, "density": 0.2}}
    print("[gpu-yarn] resampled OBJ scene, 40 steps", flush=True)
    desc, rest, glued, x = _run_scene(tmp_path, scene, 40)
    off = desc["strandOffsets"]
    assert len(off) == 3 and off[1] > 30  # resampled finer than the file
    assert desc["closed"] == [0, 1] and desc["connections"][off[1]] == off[2] - 1
