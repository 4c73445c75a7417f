"""Strain limiting after the SDF pass (kscr_set_strain_limit, following
SDF/cudaSDFCollision.cu:176-199) on the device against the FP32 restatement:
zero differing words on every path that runs the SDF pass (fused Jacobi step,
even/odd coloured step, the contact path, strands longer than a workgroup),
and the entry points that do not cover it refuse it."""
import numpy as np
import pytest

from kamisimulator_amd import scenes
from kamisimulator_amd._lib import SCHEDULE_COLORED, Batch, KscrError, Sim, default_params
from oracle.pyoracle import OracleSim
from tests import parity

pytestmark = pytest.mark.gpu
F32 = np.float32
DT = F32(1 / 600)


def _pair(sc, params, err=0.1, rest=None):
    sims = []
    for cls in (Sim, OracleSim):
        s = scenes.setup_sim(cls(params, sc.positions, sc.strand_offsets), sc)
        s.set_strain_limit(True, err, rest)
        sims.append(s)
    sims[1].set_threads(8)
    return sims


def _rest(pos, off, scale=1.0):
    rest = np.zeros(len(pos), F32)
    for s in range(len(off) - 1):
        a, b = off[s], off[s + 1]
        seg = pos[a + 1:b].astype(np.float64) - pos[a:b - 1].astype(np.float64)
        rest[a:b - 1] = (scale * np.linalg.norm(seg, axis=1)).astype(F32)
    return rest


def _run(label, g, o, steps, start):
    print(f"[gpu-strain] {label}: n={g.n} steps={steps}", flush=True)
    for _ in range(steps):
        g.solve_next_step(DT)
        o.solve_next_step(DT)
    gp, gv = g.get_pos_vel()
    op, ov = o.get_pos_vel()
    parity.check(label, {"x": gp, "v": gv, "q": g.get_q()}, {"x": op, "v": ov, "q": o.get_q()}, {"x": start})
    return gp


def _strain(pos, off, rest):
    """Largest |l - l0| / l0 over the segments."""
    worst = 0.0
    for s in range(len(off) - 1):
        a, b = off[s], off[s + 1]
        l = np.linalg.norm(pos[a + 1:b].astype(np.float64) - pos[a:b - 1], axis=1)
        worst = max(worst, float(np.max(np.abs(l - rest[a:b - 1]) / rest[a:b - 1])))
    return worst


def test_groom_pushed_out_by_the_sdf_threshold():
    """A 5 mm SDF threshold pushes the lower vertices out of the scalp and
    stretches their segments; the limit (tolerance 1e-3) resets them."""
    sc = scenes.config_groom(1000, 100, sdf_dim=64)
    sc.sdf_threshold = 5e-3
    p = scenes.make_params(default_params, sc.material)
    g, o = _pair(sc, p, err=1e-3)
    gp = _run("groom1k-threshold5mm", g, o, 20, sc.positions)
    # the same run without the limit ends elsewhere, with larger strains
    o2 = scenes.setup_sim(OracleSim(p, sc.positions, sc.strand_offsets), sc)
    o2.set_threads(8)
    for _ in range(20):
        o2.solve_next_step(DT)
    p2, _ = o2.get_pos_vel()
    rest = _rest(sc.positions, sc.strand_offsets)
    s_with, s_without = _strain(gp, sc.strand_offsets, rest), _strain(p2, sc.strand_offsets, rest)
    print(f"[gpu-strain] largest strain with the limit {s_with:.3e}, without {s_without:.3e}", flush=True)
    assert s_without > 1e-3 and s_with < s_without
    assert parity.words(gp, p2) > 0


@pytest.mark.parametrize("schedule", ["jacobi", "colored"])
def test_explicit_rest_lengths(schedule):
    """Rest lengths 10 % shorter than the groom's segments, tolerance 1e-3:
    every segment is reset after every step."""
    sc = scenes.config_groom(600, 80, sdf_dim=64)
    p = scenes.make_params(default_params, sc.material)
    if schedule == "colored":
        p.schedule = SCHEDULE_COLORED
    rest = _rest(sc.positions, sc.strand_offsets, 0.9)
    g, o = _pair(sc, p, err=1e-3, rest=rest)
    _run(f"groom600-rest0.9-{schedule}", g, o, 10, sc.positions)


def test_contact_path_and_long_strands():
    """The contact path (per-substep kernels, its own SDF kernel) on a
    lognormal groom, and strands longer than a workgroup (cut into packs)."""
    sc = scenes.config_large(400, seed=5, sdf_dim=64)
    sc.sdf_threshold = 2e-3
    p = scenes.contact_params(default_params, sc.material)
    g, o = _pair(sc, p, err=1e-3)
    _run("lognormal400-contacts", g, o, 4, sc.positions)
    # three strands of 3000, 1500 and 40 vertices hanging in a body SDF
    n = [3000, 1500, 40]
    pos, off = [], [0]
    for s, k in enumerate(n):
        z = np.arange(k, dtype=np.float64) * 1e-3
        pos.append(np.stack([np.full(k, 0.01 * s), 0.05 * np.sin(z * 3.0), -z], 1))
        off.append(off[-1] + k)
    long = scenes.Scene("long", np.concatenate(pos).astype(F32), np.array(off, np.int32))
    long.sdf = scenes.sdf_grid(scenes.sphere_sdf((0.0, -0.06, -1.5), 0.05), (-0.5, -0.5, -3.2), (0.5, 0.5, 0.2),
                               (32, 32, 96))
    long.sdf_threshold = 1e-2
    p = scenes.make_params(default_params, long.material)
    g, o = _pair(long, p, err=1e-4)
    _run("long-strands", g, o, 5, long.positions)


def test_switched_off_and_refused():
    sc = scenes.config_groom(200, 64, sdf_dim=64)
    sc.sdf_threshold = 5e-3
    p = scenes.make_params(default_params, sc.material)
    # enable = 0: the plain SDF step, word for word
    a = scenes.setup_sim(Sim(p, sc.positions, sc.strand_offsets), sc)
    b = scenes.setup_sim(Sim(p, sc.positions, sc.strand_offsets), sc)
    b.set_strain_limit(True, 1e-3)
    b.set_strain_limit(False)
    for _ in range(5):
        a.solve_next_step(DT)
        b.solve_next_step(DT)
    assert parity.words(a.get_pos_vel()[0], b.get_pos_vel()[0]) == 0
    with pytest.raises(KscrError, match="INVALID_ARGUMENT"):
        b.set_strain_limit(True, -1.0)
    with pytest.raises(KscrError, match="INVALID_ARGUMENT"):
        b.set_strain_limit(True, float("nan"))
    # the batched step and the adjoint do not cover it
    c = scenes.setup_sim(Sim(p, sc.positions, sc.strand_offsets), sc)
    c.set_strain_limit(True, 1e-3)
    bt = Batch([c])
    with pytest.raises(KscrError, match="UNSUPPORTED"):
        bt.step(DT, 1, with_sdf=True)
    bt.step(DT, 1, with_sdf=False)  # without the SDF pass the limit never runs
    with pytest.raises(KscrError, match="UNSUPPORTED"):
        c.adjoint_forward(DT, 2, 1, with_sdf=True)
