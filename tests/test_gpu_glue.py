"""Glued vertices (kscr_set_connections: Vertex::connectionIndex,
YarnBall.h:56) on the GPU against the FP32 restatement, word for word.

A groom with glued vertices steps through the per-iteration kernels of the
contact path, whose position solve adds the glue spring
(YarnBall/sim/cosserat.cu:123-130) and whose narrow phase skips glued
segment pairs (YarnBall/sim/collision.cu:44-49)."""
import numpy as np
import pytest

from kamisimulator_amd import scenes
from kamisimulator_amd._lib import SCHEDULE_COLORED, Batch, KscrError, Sim, default_params
from oracle.pyoracle import OracleSim
from tests import parity

pytestmark = pytest.mark.gpu
DT = np.float32(1 / 600)


def _rings(num=6, n=40, radius=0.01, seed=2):
    """num closed rings (last point = first point) in a row along z, each
    tilted a little; returns positions, offsets, connections, top vertices."""
    rng = np.random.default_rng(seed)
    pos, off, conn, tops = [], [0], [], []
    for r in range(num):
        th = np.linspace(0.0, 2 * np.pi, n)
        tilt = rng.uniform(-0.3, 0.3)
        p = np.stack([radius * np.cos(th), radius * np.sin(th) * np.cos(tilt), radius * np.sin(th) * np.sin(tilt)
                      + 0.003 * r], 1)
        p[-1] = p[0]
        base = off[-1]
        c = np.full(n, -1, np.int32)
        c[0], c[n - 1] = base + n - 1, base
        pos.append(p)
        conn.append(c)
        tops.append(base + int(np.argmax(p[:, 1])))
        off.append(base + n)
    return (np.concatenate(pos).astype(np.float32), np.array(off, np.int32), np.concatenate(conn),
            np.array(tops, np.int32))


def _pair(pos, off, conn, pins, params, material=scenes.Material()):
    sims = []
    for cls in (Sim, OracleSim):
        s = cls(params, pos, off)
        s.configure(material.density)
        s.set_young(material.young_gpa, material.deviation)
        s.pin(pins)
        s.set_connections(conn)
        s.upload()
        sims.append(s)
    return sims


def _run(label, g, o, steps, start):
    for _ in range(steps):
        g.solve_next_step(DT)
        o.solve_next_step(DT)
    gp, gv = g.get_pos_vel()
    op, ov = o.get_pos_vel()
    return parity.check(label, {"x": gp, "v": gv, "q": g.get_q()}, {"x": op, "v": ov, "q": o.get_q()}, start)


def test_closed_rings_under_gravity():
    pos, off, conn, tops = _rings()
    p = scenes.make_params(default_params, scenes.Material())
    g, o = _pair(pos, off, conn, tops, p)
    print(f"[gpu-glue] {len(off) - 1} closed rings, {len(pos)} vertices, 60 steps", flush=True)
    _run("closed rings", g, o, 60, {"x": pos})
    x, _ = g.get_pos_vel()
    for r in range(len(off) - 1):  # the ends stay joined
        assert np.linalg.norm(x[off[r]] - x[off[r + 1] - 1]) < 1e-6


def test_glued_endpoints_across_strands_with_contacts():
    """Strand tips glued to the middle of neighbouring strands (glueEndpoints'
    pattern), with hair-hair contacts on: the glued segment pairs are absent
    from the contact lists, and the step equals the restatement."""
    sc = scenes.ribbon(8, 24)
    n = sc.num_verts
    conn = np.full(n, -1, np.int32)
    for s in range(0, 8, 2):  # tip of strand s glued to vertex 12 of strand s + 1
        a, b = int(sc.strand_offsets[s + 1]) - 1, int(sc.strand_offsets[s + 1]) + 12
        conn[a], conn[b] = b, a
    p = scenes.contact_params(default_params, sc.material)
    g, o = _pair(sc.positions, sc.strand_offsets, conn, sc.roots, p)
    print(f"[gpu-glue] ribbon with glued tips, {n} vertices, 20 steps", flush=True)
    _run("glued ribbon with contacts", g, o, 20, {"x": sc.positions})
    gn, gc = g.get_contacts()
    on, oc = o.contacts()
    assert on.sum() > 0 and np.array_equal(gn, on) and np.array_equal(gc, oc)
    for a in range(n):  # no listed pair is exempt by the glue rule
        c = [conn[a], conn[a + 1] if a + 1 < n else -1]
        for b in gc[a, :gn[a]]:
            assert b not in c and b + 1 not in c


def test_glue_without_contacts_uses_the_iteration_kernels():
    pos, off, conn, tops = _rings(num=3)
    p = scenes.make_params(default_params, scenes.Material(wind_peak=40.0))
    g, o = _pair(pos, off, conn, tops, p)
    _run("closed rings with wind", g, o, 30, {"x": pos})


def test_connection_checks():
    pos, off, conn, tops = _rings(num=2)
    p = scenes.make_params(default_params, scenes.Material())
    g = Sim(p, pos, off)
    g.configure(1300.0)
    bad = conn.copy()
    bad[0] = 5  # not glued back
    with pytest.raises(KscrError, match="INVALID_ARGUMENT"):
        g.set_connections(bad)
    bad = conn.copy()
    bad[3] = 3
    with pytest.raises(KscrError, match="INVALID_ARGUMENT"):
        g.set_connections(bad)
    g.set_connections(conn)
    g.upload()
    pc = scenes.make_params(default_params, scenes.Material(), schedule=SCHEDULE_COLORED)
    with pytest.raises(KscrError, match="UNSUPPORTED"):
        g.set_params(pc)
    with pytest.raises(KscrError, match="UNSUPPORTED"):
        g.adjoint_forward(DT, 2, 1)
    with pytest.raises(Exception, match="UNSUPPORTED"):
        Batch([g])
    g.set_connections(None)  # unglued: every path open again
    g.upload()
    b = Batch([g])
    b.step(DT, 1)
    b.close()
