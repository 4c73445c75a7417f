"""GPU parity of the even/odd coloured schedule (kscr_colored.hip) against the
oracle's coloured iteration (scr_oracle.hpp: iterate_colored), through the C
ABI, on identical seeded inputs.  As for the Jacobi step the two evaluate the
same FP32 operations in the same order, so the agreement is bit-exact; the
asserted bar is 1e-5 relative.  Also: what the coloured schedule rejects."""
import numpy as np
import pytest

from kamisimulator_amd import scenes
from kamisimulator_amd._lib import SCHEDULE_COLORED, Batch, KscrError, Sim, default_params
from oracle.pyoracle import OracleSim
from tests import parity

pytestmark = pytest.mark.gpu
F32 = np.float32
DT = F32(1 / 600)
TOL = 1e-5


def _rel(a, b):
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max() / max(np.abs(b).max(), 1e-30))


def _run(sc, steps, block=0, label=""):
    print(f"[gpu-colored] {label} verts={sc.num_verts} strands={sc.num_strands} steps={steps} block={block}", flush=True)
    p = scenes.make_params(default_params, sc.material, block_size=block, schedule=SCHEDULE_COLORED)
    g = scenes.setup_sim(Sim(p, sc.positions, sc.strand_offsets), sc)
    o = scenes.setup_sim(OracleSim(p, sc.positions, sc.strand_offsets, precision=32), sc)
    o.set_threads(8)
    q0 = g.get_q()
    for _ in range(steps):
        g.solve_next_step(DT)
        o.solve_next_step(DT)
    gp, gv = g.get_pos_vel()
    op, ov = o.get_pos_vel()
    gq, oq = g.get_q(), o.get_q()
    print(f"[gpu-colored] {label}: block {g.block_size}", flush=True)
    r = parity.check(f"colored {label}", {"x": gp, "v": gv, "q": gq}, {"x": op, "v": ov, "q": oq},
                     {"x": sc.positions, "q": q0})
    ep, eq, ev = r["x"]["scale_rel"], r["q"]["scale_rel"], r["v"]["scale_rel"]
    return ep, eq, ev, gp


def test_colored_cantilever_100_steps():
    ep, eq, ev, _ = _run(scenes.config_cantilever(), 100, label="cantilever")
    assert ep <= TOL and eq <= TOL and ev <= TOL


def test_colored_groom_with_sdf_and_wind():
    sc = scenes.config_groom(1000, 100, with_sdf=True, sdf_dim=64)
    sc.material = scenes.Material(wind_peak=40.0)
    ep, eq, ev, _ = _run(sc, 50, label="groom1k-sdf-wind")
    assert ep <= TOL and eq <= TOL and ev <= TOL


def test_colored_lognormal_lengths():
    # strands of 32..256 vertices start at odd and even global indices
    sc = scenes.config_large(600, seed=5, sdf_dim=64)
    assert (sc.strand_offsets[:-1] % 2 == 1).any() and (sc.strand_offsets[:-1] % 2 == 0).any()
    ep, eq, ev, _ = _run(sc, 30, label="lognormal600")
    assert ep <= TOL and eq <= TOL and ev <= TOL


@pytest.mark.parametrize("block", [128, 256, 512])
def test_colored_block_sizes_agree(block):
    sc = scenes.config_groom(200, 100, with_sdf=False)
    ep, eq, ev, _ = _run(sc, 20, block=block, label=f"block{block}")
    assert ep == 0.0 and eq == 0.0 and ev == 0.0


def test_colored_differs_from_jacobi():
    sc = scenes.config_groom(100, 100, with_sdf=False)
    p = scenes.make_params(default_params, sc.material)
    j = scenes.setup_sim(Sim(p, sc.positions, sc.strand_offsets), sc)
    j.solve_next_step(DT)
    _, _, _, cp = _run(sc, 1, label="one-step")
    jp, _ = j.get_pos_vel()
    assert np.abs(cp - jp).max() > 0


def test_colored_rejections():
    sc = scenes.config_groom(50, 40, with_sdf=False)
    p = scenes.make_params(default_params, sc.material, schedule=SCHEDULE_COLORED)
    p.enable_hair_hair_coll = 1
    with pytest.raises(KscrError):
        scenes.setup_sim(Sim(p, sc.positions, sc.strand_offsets), sc)
    p = scenes.make_params(default_params, sc.material, schedule=SCHEDULE_COLORED)
    g = scenes.setup_sim(Sim(p, sc.positions, sc.strand_offsets), sc)
    with pytest.raises(KscrError):
        g.adjoint_forward(DT, 2, 1)
    with pytest.raises(KscrError):
        Batch([g])
    # a strand longer than the coloured kernel's largest pack (1023 vertices)
    n = 1100
    pos = np.stack([np.arange(n) * 1e-3, np.zeros(n), np.zeros(n)], 1).astype(F32)
    long_sc = scenes.Scene("long", pos, np.array([0, n], np.int32))
    with pytest.raises(KscrError):
        scenes.setup_sim(Sim(p, long_sc.positions, long_sc.strand_offsets), long_sc)
