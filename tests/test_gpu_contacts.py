"""The step with hair-hair contacts (LBVH broad phase, narrow phase, IPC
barrier and friction, step limits) against the FP32 restatement."""
import numpy as np
import pytest

from kamisimulator_amd import scenes
from kamisimulator_amd._lib import KscrError, Sim, default_params
from oracle.pyoracle import OracleSim
from tests import parity

pytestmark = pytest.mark.gpu
DT = np.float32(1 / 600)


def _pair(sc, **kw):
    p = scenes.contact_params(default_params, sc.material, **kw)
    g = scenes.setup_sim(Sim(p, sc.positions, sc.strand_offsets), sc)
    o = scenes.setup_sim(OracleSim(p, sc.positions, sc.strand_offsets), sc)
    o.set_threads(8)
    return g, o


def _rel(a, b):
    return float(np.abs(a.astype(np.float64) - b).max() / np.abs(b).max())


def test_ribbon_contacts_match_oracle():
    sc = scenes.ribbon(12, 32)
    rng = np.random.default_rng(3)
    vel = rng.uniform(-0.02, 0.02, size=sc.positions.shape).astype(np.float32)
    g, o = _pair(sc)
    g.set_pos_vel(sc.positions, vel)
    o.set_pos_vel(sc.positions, vel)
    print(f"[gpu-contacts] bundle verts={sc.num_verts}", flush=True)
    for s in range(10):
        g.solve_next_step(DT)
        o.solve_next_step(DT)
    gp, gv = g.get_pos_vel()  # raises if the device flagged an overflow
    assert o.error() == 0
    gn, gc = g.get_contacts()
    on, oc = o.contacts()
    print(f"[gpu-contacts] contacts gpu {gn.sum()} oracle {on.sum()} pairs(oracle) {o.num_pairs()}", flush=True)
    assert on.sum() > 50
    assert np.array_equal(gn, on) and np.array_equal(gc, oc)
    op, ov = o.get_pos_vel()
    parity.check("ribbon-10-steps", {"x": gp, "v": gv, "q": g.get_q()}, {"x": op, "v": ov, "q": o.get_q()},
                 {"x": sc.positions})


def test_groom_with_contacts_and_sdf():
    counts = scenes.lognormal_counts(300, seed=8)
    sc = scenes.groom(300, counts, seed=8, sphere_radius=0.02)  # dense scalp -> contacts
    sc.sdf = scenes.sdf_grid(scenes.sphere_sdf((0, 0, 0), 0.02), (-0.1,) * 3, (0.1,) * 3, (64,) * 3)
    g, o = _pair(sc)
    for s in range(5):
        g.solve_next_step(DT)
        o.solve_next_step(DT)
    gn, _ = g.get_contacts()
    on, _ = o.contacts()
    print(f"[gpu-contacts] groom300 contacts gpu {gn.sum()} oracle {on.sum()}", flush=True)
    assert np.array_equal(gn, on)
    gp, gv = g.get_pos_vel()
    op, ov = o.get_pos_vel()
    parity.check("groom300-contacts-sdf", {"x": gp, "v": gv, "q": g.get_q()}, {"x": op, "v": ov, "q": o.get_q()},
                 {"x": sc.positions})


def test_dense_lognormal_groom_with_contacts():
    """A lognormal groom (config 2's generator at 500 strands) with contacts and
    a coarse sphere + capsule SDF: interpenetrating strands hit the skipped
    non-finite steps; GPU and restatement must still agree word for word."""
    sc = scenes.config_large(500, seed=2, sdf_dim=64)
    g, o = _pair(sc)
    print(f"[gpu-contacts] lognormal500 verts={sc.num_verts}", flush=True)
    for _ in range(3):
        g.solve_next_step(DT)
        o.solve_next_step(DT)
    gp, gv = g.get_pos_vel()
    op, ov = o.get_pos_vel()
    gn, _ = g.get_contacts()
    on, _ = o.contacts()
    nbit = int(np.count_nonzero(gp.view(np.uint32) != op.view(np.uint32)))
    print(f"[gpu-contacts] lognormal500: contacts gpu {gn.sum()} oracle {on.sum()} differing pos words {nbit}", flush=True)
    assert np.isfinite(gp).all() and np.array_equal(gn, on)
    parity.check("lognormal500-contacts", {"x": gp, "v": gv, "q": g.get_q()}, {"x": op, "v": ov, "q": o.get_q()},
                 {"x": sc.positions})


def test_too_many_contacts_is_reported():
    # 30 strands packed within one barrier width: > 16 contacts per segment
    sc = scenes.bundle(6, 5, 16, spacing=1.5e-5, jitter=2e-6)
    p = scenes.contact_params(default_params, sc.material)
    g = scenes.setup_sim(Sim(p, sc.positions, sc.strand_offsets), sc)
    g.solve_next_step(DT)
    with pytest.raises(KscrError) as ei:
        g.get_pos_vel()
    assert "MAX_COLLISIONS" in str(ei.value)
    o = scenes.setup_sim(OracleSim(p, sc.positions, sc.strand_offsets), sc)
    o.solve_next_step(DT)
    assert o.error() == 8


def _packs(offsets, capacity):
    """Greedy packs of whole strands (the library's packing)."""
    packs, s, ns = [], 0, len(offsets) - 1
    while s < ns:
        beg = end = offsets[s]
        while s < ns and offsets[s + 1] - beg <= capacity:
            end = offsets[s + 1]
            s += 1
        packs.append((beg, end))
    return packs


@pytest.mark.parametrize("num_itr", [10, 7])
def test_packs_without_contacts_match_oracle_bitwise(num_itr):
    """Packs with no contact in a substep run all iterations in one launch
    (neighbour iterates in LDS), the others one launch per iteration.  A groom
    where both kinds occur, with an even and an odd iteration count (the two
    ends of the dx / q ping-pong), must equal the FP32 restatement word for
    word."""
    sc = scenes.config_large(400, seed=5, sdf_dim=64)
    sc.material.num_itr = num_itr
    g, o = _pair(sc)
    packs = _packs(sc.strand_offsets, g.block_size)
    mixed = 0
    for _ in range(4):
        g.solve_next_step(DT)
        o.solve_next_step(DT)
        gn, gc = g.get_contacts()
        hit = np.array([bool((gn[b:e] > 0).any()) for b, e in packs])
        mixed += int(hit.any() and not hit.all())
        print(f"[gpu-contacts] itr {num_itr}: packs with contacts {int(hit.sum())} of {len(hit)}", flush=True)
    gp, gv = g.get_pos_vel()
    op, ov = o.get_pos_vel()
    on, oc = o.contacts()
    assert mixed == 4
    assert np.array_equal(gn, on) and np.array_equal(gc, oc)
    assert np.array_equal(gp.view(np.uint32), op.view(np.uint32))
    assert np.array_equal(gv.view(np.uint32), ov.view(np.uint32))
    assert np.array_equal(g.get_q().view(np.uint32), o.get_q().view(np.uint32))


def test_periodic_bvh_rebuild_70_steps():
    """70 steps = 350 substeps of 1/3000 s: the tree is built at the first
    detection and rebuilt once lastBVHRebuild reaches 0.1 s (300 substeps;
    YBSimulator.cu:436-444), with refits in between.  Word for word against the
    FP32 restatement."""
    sc = scenes.ribbon(12, 32)
    rng = np.random.default_rng(3)
    vel = rng.uniform(-0.02, 0.02, size=sc.positions.shape).astype(np.float32)
    g, o = _pair(sc)
    g.set_pos_vel(sc.positions, vel)
    o.set_pos_vel(sc.positions, vel)
    q0 = g.get_q()
    print(f"[gpu-contacts] ribbon verts={sc.num_verts} steps=70", flush=True)
    for s in range(70):
        g.solve_next_step(DT)
        o.solve_next_step(DT)
    gn, gc = g.get_contacts()
    on, oc = o.contacts()
    assert on.sum() > 0 and np.array_equal(gn, on) and np.array_equal(gc, oc)
    gp, gv = g.get_pos_vel()
    op, ov = o.get_pos_vel()
    parity.check("ribbon-70-steps", {"x": gp, "v": gv, "q": g.get_q()}, {"x": op, "v": ov, "q": o.get_q()},
                 {"x": sc.positions, "q": q0})


def test_large_groom_with_contacts_full_size():
    """The large groom at its full size: 100k strands with lognormal vertex
    counts (median 96, clipped 32-256, 10.2M vertices), sphere + capsule SDF on
    256^3, hair-hair contacts, 100 steps of 1/600 s (500 substeps, so the
    0.1 s tree rebuild runs too).  Roots on a spiral and 0.2-0.6 mm curls:
    uniformly random roots put some roots microns apart, and the strands that
    grow from them run side by side with more than the 16 contacts per segment
    the reference treats as fatal.

    The FP32 restatement is too slow at this size for the whole rollout, so the
    checks are size-independent:
    - finite state, unit rotors, pinned roots unmoved;
    - no device error (16-contact lists, pair buffer, query stack);
    - one more substep's detection from the device's own state (x, v, the
      previous substep's velocities, q) gives contact lists equal, word for
      word, to the restatement's LBVH broad phase plus narrow phase on that
      same state."""
    sc = scenes.config_large(100_000, seed=2, sdf_dim=256, curl=(2e-4, 6e-4), roots="spiral")
    p = scenes.contact_params(default_params, sc.material)
    print(f"[gpu-contacts] large groom verts={sc.num_verts} strands={sc.num_strands} steps=100", flush=True)
    g = scenes.setup_sim(Sim(p, sc.positions, sc.strand_offsets), sc)
    for s in range(100):
        g.solve_next_step(DT)
        if (s + 1) % 20 == 0:
            g.synchronize()
            print(f"[gpu-contacts] large groom: step {s + 1}", flush=True)
    x, v = g.get_pos_vel()  # raises on a device error
    u, q = g.get_last_vel(), g.get_q()
    n, _ = g.get_contacts()
    print(f"[gpu-contacts] large groom: contacts after 100 steps {int(n.sum())}, max per segment {int(n.max())}",
          flush=True)
    assert np.isfinite(x).all() and np.isfinite(v).all() and np.isfinite(q).all()
    assert np.array_equal(x[sc.roots].view(np.uint32), sc.positions[sc.roots].view(np.uint32))
    np.testing.assert_allclose(np.linalg.norm(q.astype(np.float64), axis=1), 1.0, atol=1e-5)
    assert n.sum() > 1000
    # one substep (h = 1/3000 s, the step's own substep length, so the clock
    # and the adaptive guess see the same h and last h): detection first
    sub_dt = np.float32(DT / np.float32(5))
    g.advance(sub_dt)
    gn, gc = g.get_contacts()
    g.get_pos_vel()
    o = scenes.setup_sim(OracleSim(p, sc.positions, sc.strand_offsets), sc)
    o.set_threads(16)
    for rep in range(2):  # the first call sets the oracle's substep length
        o.set_pos_vel(x, v)
        o.set_last_vel(u)
        o.set_q(q)
        o.detect_only(sub_dt)
    on, oc = o.contacts()
    print(f"[gpu-contacts] large groom detection: contacts gpu {int(gn.sum())} oracle {int(on.sum())}, "
          f"broad-phase pairs {o.num_pairs()}", flush=True)
    assert o.error() == 0
    assert np.array_equal(gn, on) and np.array_equal(gc, oc)
