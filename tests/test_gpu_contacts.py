"""The step with hair-hair contacts (LBVH broad phase, narrow phase, IPC
barrier and friction, step limits) against the FP32 restatement."""
import numpy as np
import pytest

from kamisimulator_amd import scenes
from kamisimulator_amd._lib import KscrError, Sim, default_params
from oracle.pyoracle import OracleSim

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
    ep, eq = _rel(gp, op), _rel(g.get_q(), o.get_q())
    print(f"[gpu-contacts] pos rel {ep:.3e} q rel {eq:.3e} differing pos words "
          f"{int(np.count_nonzero(gp.view(np.uint32) != op.view(np.uint32)))}", flush=True)
    assert ep <= 1e-5 and eq <= 1e-5


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
    assert _rel(g.get_pos_vel()[0], o.get_pos_vel()[0]) <= 1e-5


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
    assert _rel(gp, op) <= 1e-5 and _rel(gv, ov) <= 1e-5


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
