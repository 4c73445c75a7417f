"""Batched form (kscr_batch_*) and the autotuner on the GPU.

A batch advances several independent grooms (own sizes, gravity directions
and SDF grids) with one launch per step replayed from a hipGraph, under
either mapping of packs onto workgroups (one per pack, or a persistent grid);
each groom must match the FP32 restatement run on its own word for word."""
import numpy as np
import pytest

from kamisimulator_amd import scenes
from kamisimulator_amd._lib import BATCH_GRID, BATCH_PERSISTENT, Batch, Sim, default_params
from oracle.pyoracle import OracleSim
from tests import parity

pytestmark = pytest.mark.gpu
DT = np.float32(1 / 600)


def _grooms(n=4, seed=11):
    rng = np.random.default_rng(seed)
    out = []
    for g in range(n):
        ns = int(rng.integers(50, 300))
        counts = rng.integers(32, 257, size=ns)
        sc = scenes.groom(ns, counts, seed=seed + g, center=(0.0, 0.0, 0.0))
        R = 0.1
        sc.sdf = scenes.sdf_grid(scenes.sphere_sdf((0, 0, 0), R), (-1.5 * R,) * 3, (1.5 * R,) * 3, (48,) * 3)
        d = np.array([rng.normal() * 0.3, -1.0, rng.normal() * 0.3])
        d /= np.linalg.norm(d)
        sc.material = scenes.Material(gravity=tuple(9.81 * d))
        out.append(sc)
    return out


def _make(sc, cls, **kw):
    p = scenes.make_params(default_params, sc.material)
    return scenes.setup_sim(cls(p, sc.positions, sc.strand_offsets, **kw), sc)


def _rel(a, b):
    return float(np.abs(a.astype(np.float64) - b).max() / np.abs(b).max())


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("mapping", [BATCH_GRID, BATCH_PERSISTENT])
def test_batch_matches_oracle(use_graph, mapping):
    grooms = _grooms()
    print(f"[gpu-batch] grooms={len(grooms)} verts={[g.num_verts for g in grooms]} graph={use_graph} "
          f"mapping={mapping}", flush=True)
    sims = [_make(sc, Sim) for sc in grooms]
    orcs = [_make(sc, OracleSim) for sc in grooms]
    for o in orcs:
        o.set_threads(8)
    b = Batch(sims)
    b.set_graph(use_graph)
    b.set_mapping(mapping)
    assert b.mapping == mapping
    b.step(DT, nsteps=4, with_sdf=True)
    b.step(DT, nsteps=4, with_sdf=True)
    b.synchronize()
    for o in orcs:
        for _ in range(8):
            o.solve_next_step(DT)
    for g, (s, o, sc) in enumerate(zip(sims, orcs, grooms)):
        gp, gv = s.get_pos_vel()
        op, ov = o.get_pos_vel()
        parity.check(f"batch groom {g}", {"x": gp, "v": gv, "q": s.get_q()}, {"x": op, "v": ov, "q": o.get_q()},
                     {"x": sc.positions})


@pytest.mark.parametrize("mapping", [BATCH_GRID, BATCH_PERSISTENT])
def test_batch_equals_per_groom_launches(mapping):
    grooms = _grooms(3, seed=5)
    a = [_make(sc, Sim) for sc in grooms]
    c = [_make(sc, Sim) for sc in grooms]
    b = Batch(a)
    b.set_mapping(mapping)
    b.step(DT, nsteps=6, with_sdf=True)
    b.synchronize()
    for s in c:
        for _ in range(6):
            s.solve_next_step(DT)
    for s, t in zip(a, c):
        assert np.array_equal(s.get_pos_vel()[0].view(np.uint32), t.get_pos_vel()[0].view(np.uint32))
        # a handle stepped by a batch keeps stepping on its own afterwards
    for s, t in zip(a, c):
        s.solve_next_step(DT)
        t.solve_next_step(DT)
        assert np.array_equal(s.get_q().view(np.uint32), t.get_q().view(np.uint32))


def test_autotune_keeps_state():
    sc = scenes.config_groom(2000, 100, sdf_dim=64)
    a, c = _make(sc, Sim), _make(sc, Sim)
    best = a.autotune(DT, steps=2)
    print(f"[gpu-batch] autotune picked block {best}", flush=True)
    assert best in (128, 256, 512, 1024)
    for s in (a, c):
        for _ in range(3):
            s.solve_next_step(DT)
    assert np.array_equal(a.get_pos_vel()[0].view(np.uint32), c.get_pos_vel()[0].view(np.uint32))


def test_batch_autotune_picks_size_and_mapping():
    """The batch autotuner times every workgroup size under both mappings and
    keeps the fastest pair; the grooms' states and clocks come back as they
    were, so the batch then steps exactly as per-groom launches do."""
    grooms = _grooms(3, seed=8)
    a = [_make(sc, Sim) for sc in grooms]
    c = [_make(sc, Sim) for sc in grooms]
    b = Batch(a)
    best = b.autotune(DT, steps=2)
    print(f"[gpu-batch] batch autotune: block {best}, mapping {b.mapping}", flush=True)
    assert best in (256, 512, 1024) and b.mapping in (BATCH_GRID, BATCH_PERSISTENT)
    with pytest.raises(Exception):
        b.set_mapping(7)
    b.step(DT, nsteps=3, with_sdf=True)
    b.synchronize()
    for s in c:
        for _ in range(3):
            s.solve_next_step(DT)
    for s, t in zip(a, c):
        assert np.array_equal(s.get_pos_vel()[0].view(np.uint32), t.get_pos_vel()[0].view(np.uint32))


def test_64_grooms_200_steps_full_size():
    """The batched workload at its full size: 64 grooms with log-uniform
    1k-50k strands of 32-256 vertices (about 9e7 vertices), each with its own
    128^3 sphere SDF and gravity direction, 200 steps in one graph-replayed
    launch per step.  Grooms and strands are independent, so each groom's
    first and last 3 strands, run alone in the FP32 restatement with that
    groom's gravity and SDF, must equal the batch word for word."""
    grooms = scenes.config_batched()
    nv = sum(sc.num_verts for sc in grooms)
    print(f"[gpu-batch] 64 grooms, {nv} vertices, 200 steps", flush=True)
    sims = [_make(sc, Sim) for sc in grooms]
    b = Batch(sims)
    b.set_graph(True)
    for k in range(4):
        # the first 100 steps one workgroup per pack, the last 100 persistent
        b.set_mapping(BATCH_GRID if k < 2 else BATCH_PERSISTENT)
        b.step(DT, nsteps=50, with_sdf=True)
        b.synchronize()
        print(f"[gpu-batch] step {50 * (k + 1)}", flush=True)
    b.close()
    for gi, (sc, s) in enumerate(zip(grooms, sims)):
        x, v = s.get_pos_vel()
        q = s.get_q()
        assert np.isfinite(x).all() and np.isfinite(q).all()
        ns = sc.num_strands
        for s0, s1 in ((0, 3), (ns - 3, ns)):
            a, e = int(sc.strand_offsets[s0]), int(sc.strand_offsets[s1])
            part = scenes.Scene(f"batched{gi}-{s0}", sc.positions[a:e].copy(),
                                (sc.strand_offsets[s0:s1 + 1] - a).astype(np.int32), material=sc.material)
            part.sdf = sc.sdf
            o = _make(part, OracleSim)
            q0 = o.get_q()
            for _ in range(200):
                o.solve_next_step(DT)
            op, ov = o.get_pos_vel()
            parity.check(f"groom {gi} strands {s0}-{s1}", {"x": x[a:e], "v": v[a:e], "q": q[a:e]},
                         {"x": op, "v": ov, "q": o.get_q()}, {"x": part.positions, "q": q0})
        s.close()
