"""Strands longer than the largest workgroup (1024 vertices) on the GPU
against the FP32 restatement, word for word.

The reference runs chains of any length in overlapping blocks with a
one-vertex halo (YarnBall/sim/cosserat.cu:30-37).  Here such a strand is cut
into packs that overlap by one vertex at each cut and steps through the
per-iteration kernels (kscr_contact.hip: iterate_kernel); the restatement
has no packs at all, so equal words show that the cuts change nothing."""
import numpy as np
import pytest

from kamisimulator_amd import scenes
from kamisimulator_amd._lib import SCHEDULE_COLORED, Batch, KscrError, Sim, default_params
from oracle.pyoracle import OracleSim
from tests import parity

pytestmark = pytest.mark.gpu
DT = np.float32(1 / 600)


def _curl(n, seg=1e-3, radius=2e-3, drop=3e-4, start=(0.0, 0.0, 0.0), phase=0.0):
    """A helix hanging down from `start`: n vertices `seg` apart, `drop` lower
    per vertex."""
    dth = 2 * np.arcsin(np.sqrt(seg * seg - drop * drop) / (2 * radius))
    th = phase + dth * np.arange(n)
    p = np.stack([radius * np.cos(th), -drop * np.arange(n), radius * np.sin(th)], 1)
    return (p - p[0] + np.asarray(start)).astype(np.float32)


def _groom(lengths, spacing=6e-3):
    pos, off = [], [0]
    for s, n in enumerate(lengths):
        pos.append(_curl(n, start=(spacing * s, 0.0, 0.0), phase=0.7 * s))
        off.append(off[-1] + n)
    return np.concatenate(pos), np.array(off, np.int32)


def _pair(pos, off, params, pins=None, conn=None, material=scenes.Material(), sdf=None):
    sims = []
    for cls in (Sim, OracleSim):
        s = cls(params, pos, off)
        s.configure(material.density)
        s.set_young(material.young_gpa, material.deviation)
        s.pin(off[:-1] if pins is None else pins)
        if conn is not None:
            s.set_connections(conn)
        s.upload()
        if sdf is not None:
            s.set_sdf(sdf, 0.0)
        sims.append(s)
    sims[1].set_threads(8)
    return sims


def _run(label, g, o, steps, start, with_sdf=False):
    for _ in range(steps):
        if with_sdf:
            g.solve_next_step(DT)
            o.solve_next_step(DT)
        else:
            g.advance(DT)
            o.advance(DT)
    gp, gv = g.get_pos_vel()
    op, ov = o.get_pos_vel()
    return parity.check(label, {"x": gp, "v": gv, "q": g.get_q()}, {"x": op, "v": ov, "q": o.get_q()}, start)


def test_3000_vertex_strand_with_short_ones():
    pos, off = _groom([3000, 40, 1025, 100, 1024])
    p = scenes.make_params(default_params, scenes.Material(num_itr=10))
    g, o = _pair(pos, off, p)
    assert g.block_size == 256  # long strands run in 256-vertex packs, cut with halos
    print(f"[gpu-long] strands {np.diff(off).tolist()}, 40 steps", flush=True)
    _run("3000-vertex strand", g, o, 40, {"x": pos})
    x, _ = g.get_pos_vel()
    assert np.isfinite(x).all()
    np.testing.assert_array_equal(x[off[:-1]], pos[off[:-1]])
    # the state moved (the parity check compares against the displacement too)
    assert np.abs(x - pos).max() > 1e-6


def test_cuts_at_pinned_vertices_and_small_packs():
    """Workgroups of 64 (cuts every 62 vertices), vertices pinned on both
    sides of cuts and on halos, wind and a body SDF."""
    pos, off = _groom([1500, 2100])
    pins = sorted(set(off[:-1].tolist()) | {62, 63, 64, 123, 124, 700, 1500 + 62, 1500 + 1000})
    m = scenes.Material(num_itr=8, wind_peak=40.0)
    p = scenes.make_params(default_params, m, block_size=64)
    R = 0.05
    grid = scenes.sdf_grid(scenes.sphere_sdf((0.0, -0.06, 0.0), R), (-0.1, -0.15, -0.1), (0.1, 0.05, 0.1), (48,) * 3)
    g, o = _pair(pos, off, p, pins=pins, material=m, sdf=grid)
    assert g.block_size == 64
    print("[gpu-long] 64-vertex packs, pins at cuts, wind + SDF, 30 steps", flush=True)
    _run("cuts at pins, wind, SDF", g, o, 30, {"x": pos}, with_sdf=True)
    x, _ = g.get_pos_vel()
    # pinned vertices stay, except where the SDF pass pushes them out of the body
    outside = [k for k in pins if np.linalg.norm(pos[k] - np.array([0.0, -0.06, 0.0])) > R + 1e-3]
    assert len(outside) >= 3
    np.testing.assert_array_equal(x[outside], pos[outside])


def test_long_strands_with_contacts():
    """A ribbon of four 1600-vertex strands 60 um apart: hair-hair contacts
    between cut packs."""
    sc = scenes.ribbon(4, 1600)
    m = scenes.Material(num_itr=6)
    p = scenes.contact_params(default_params, m)
    g, o = _pair(sc.positions, sc.strand_offsets, p, pins=sc.roots, material=m)
    print("[gpu-long] ribbon of four 1600-vertex strands with contacts, 10 steps", flush=True)
    _run("long strands with contacts", g, o, 10, {"x": sc.positions})
    gn, gc = g.get_contacts()
    on, oc = o.contacts()
    print(f"[gpu-long] contact entries {int(on.sum())}", flush=True)
    assert on.sum() > 0 and np.array_equal(gn, on) and np.array_equal(gc, oc)


def test_long_closed_ring():
    n = 2000
    th = np.linspace(0.0, 2 * np.pi, n)
    r = n * 1e-3 / (2 * np.pi)
    pos = np.stack([r * np.cos(th), r * np.sin(th), 0.02 * np.sin(3 * th)], 1).astype(np.float32)
    pos[-1] = pos[0]
    off = np.array([0, n], np.int32)
    conn = np.full(n, -1, np.int32)
    conn[0], conn[-1] = n - 1, 0
    top = int(np.argmax(pos[:, 1]))
    p = scenes.make_params(default_params, scenes.Material())
    g, o = _pair(pos, off, p, pins=[top], conn=conn)
    print("[gpu-long] closed 2000-vertex ring, 30 steps", flush=True)
    _run("long closed ring", g, o, 30, {"x": pos})


def test_long_strand_limits():
    pos, off = _groom([1300, 50])
    p = scenes.make_params(default_params, scenes.Material())
    g = Sim(p, pos, off)
    g.configure(1300.0)
    g.pin(off[:-1])
    g.upload()
    with pytest.raises(KscrError, match="UNSUPPORTED"):
        g.adjoint_forward(DT, 2, 1)
    with pytest.raises(Exception, match="UNSUPPORTED"):
        Batch([g])
    # the coloured schedule holds at most 1023 vertices in a workgroup
    pc = scenes.make_params(default_params, scenes.Material(), schedule=SCHEDULE_COLORED)
    with pytest.raises(KscrError, match="STRAND_TOO_LONG"):
        g.set_params(pc)
        g.upload()
    g.close()
