"""GPU parity: the HIP step (libkscr.so through the C ABI) against the FP32
CPU restatement (oracle) on identical seeded inputs.

Both sides evaluate the same FP32 operations in the same order with IEEE
division/square root and no FMA contraction, so the expected agreement is
bit-exact; the asserted bar is the 1e-5 relative tolerance
max|x - x_ref| / max|x_ref| for positions and orientations.  The FP64 error of
the FP32 result is printed per workload (it is rounding-noise dominated: the
stretch stiffness is ~1e5 x the inertia)."""
import numpy as np
import pytest

from kamisimulator_amd import scenes
from kamisimulator_amd._lib import KscrError
from kamisimulator_amd._lib import Sim, default_params
from oracle.pyoracle import OracleSim
from tests import parity

pytestmark = pytest.mark.gpu
F32 = np.float32
DT = F32(1 / 600)
TOL = 1e-5


def _pair(sc, block=0, fp64=False):
    p = scenes.make_params(default_params, sc.material, block_size=block)
    g = scenes.setup_sim(Sim(p, sc.positions, sc.strand_offsets), sc)
    o = scenes.setup_sim(OracleSim(p, sc.positions, sc.strand_offsets, precision=32), sc)
    o.set_threads(8)
    o64 = None
    if fp64:
        o64 = scenes.setup_sim(OracleSim(p, sc.positions, sc.strand_offsets, precision=64), sc)
        o64.set_threads(8)
    return g, o, o64


def _rel(a, b):
    return parity.scale_rel(a, b)


def _run(sc, steps, block=0, fp64=False, label="", threads=8):
    print(f"[gpu-parity] {label} verts={sc.num_verts} strands={sc.num_strands} steps={steps} block={block}", flush=True)
    g, o, o64 = _pair(sc, block, fp64)
    o.set_threads(threads)
    start = {"x": sc.positions, "q": g.get_q()}
    for s in range(steps):
        g.solve_next_step(DT)
        o.solve_next_step(DT)
        if o64 is not None:
            o64.solve_next_step(DT)
        if steps >= 50 and (s + 1) % 25 == 0:
            print(f"[gpu-parity] {label}: step {s + 1}", flush=True)
    gp, gv = g.get_pos_vel()
    op, ov = o.get_pos_vel()
    gq, oq = g.get_q(), o.get_q()
    r = parity.check(label, {"x": gp, "v": gv, "q": gq}, {"x": op, "v": ov, "q": oq}, start)
    if o64 is not None:
        p64, _ = o64.get_pos_vel(f64=True)
        print(f"[gpu-parity] {label}: FP32 vs FP64 pos: rel {parity.scale_rel(gp, p64):.3e}, "
              f"disp-rel {parity.disp_rel(gp, p64, sc.positions):.3e}", flush=True)
    return r["x"]["scale_rel"], r["q"]["scale_rel"], r["v"]["scale_rel"], g, o


def test_download_views_equal_copies():
    sc = scenes.config_groom(200, 64, with_sdf=False)
    g, _, _ = _pair(sc)
    for _ in range(3):
        g.solve_next_step(DT)
    p, v = g.get_pos_vel()
    pv, vv = g.download_pos_vel()
    assert np.array_equal(p.view(np.uint32), pv.view(np.uint32))
    assert np.array_equal(v.view(np.uint32), vv.view(np.uint32))
    assert not pv.flags.writeable


def test_setup_matches_oracle():
    sc = scenes.config_groom(300, 100, with_sdf=False)
    g, o, _ = _pair(sc)
    rg, ro = g.get_rest(), o.get_rest()
    for k in rg:
        assert np.array_equal(rg[k].view(np.uint32), ro[k].view(np.uint32)), k
    assert np.array_equal(g.get_q().view(np.uint32), o.get_q().view(np.uint32))


def test_cantilever_100_steps():
    ep, eq, ev, _, _ = _run(scenes.config_cantilever(), 100, fp64=True, label="cantilever")
    assert ep <= TOL and eq <= TOL


def test_groom_with_wind_100_steps():
    """Wind on (the SCR solver's default WindPeak 40, sharpness 3): cos and
    pow of initItr come from the deterministic sequences both sides share."""
    sc = scenes.config_groom(1000, 100, with_sdf=True, sdf_dim=64)
    sc.material = scenes.Material(wind_peak=40.0)
    ep, eq, ev, _, _ = _run(sc, 100, fp64=True, label="groom1k-wind")
    assert ep <= TOL and eq <= TOL and ev <= TOL


def test_cantilever_root_pin_only():
    sc = scenes.cantilever(64, 1e-3, clamp=False)
    ep, eq, _, _, _ = _run(sc, 100, label="cantilever-pin-only")
    assert ep <= TOL and eq <= TOL


def test_groom_with_sdf_100_steps():
    sc = scenes.config_groom(1000, 100, sdf_dim=128)
    ep, eq, _, _, _ = _run(sc, 100, fp64=True, label="groom1k-sdf128")
    assert ep <= TOL and eq <= TOL


def test_lognormal_lengths_with_sdf():
    counts = scenes.lognormal_counts(600, seed=5)
    sc = scenes.groom(600, counts, seed=5)
    sc.sdf = scenes.config_large(1, with_sdf=True, sdf_dim=96).sdf
    ep, eq, _, _, _ = _run(sc, 30, label="lognormal600-sdf96")
    assert ep <= TOL and eq <= TOL


@pytest.mark.parametrize("block", [256, 512, 1024])
def test_block_sizes_agree(block):
    sc = scenes.config_groom(400, 100, sdf_dim=64)
    ep, eq, _, _, _ = _run(sc, 10, block=block, label=f"block{block}")
    assert ep <= TOL and eq <= TOL


def test_run_to_run_deterministic():
    sc = scenes.config_groom(500, 100, sdf_dim=64)
    p = scenes.make_params(default_params, sc.material)
    outs = []
    for _ in range(2):
        g = scenes.setup_sim(Sim(p, sc.positions, sc.strand_offsets), sc)
        for s in range(20):
            g.solve_next_step(DT)
        outs.append(np.concatenate([g.get_pos_vel()[0].ravel(), g.get_q().ravel()]))
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


def test_advance_without_sdf_matches():
    sc = scenes.config_groom(200, 100, sdf_dim=64)
    p = scenes.make_params(default_params, sc.material)
    g = scenes.setup_sim(Sim(p, sc.positions, sc.strand_offsets), sc)
    o = scenes.setup_sim(OracleSim(p, sc.positions, sc.strand_offsets), sc)
    for s in range(10):
        g.advance(DT)
        o.advance(DT)
    gp, gv = g.get_pos_vel()
    op, ov = o.get_pos_vel()
    parity.check("advance-no-sdf", {"x": gp, "v": gv, "q": g.get_q()}, {"x": op, "v": ov, "q": o.get_q()},
                 {"x": sc.positions})


def test_headline_size_properties():
    """100k strands x 100 vertices (the headline workload): size-independent
    properties after 5 steps, since the oracle is too slow at this size."""
    sc = scenes.config_groom(100_000, 100, with_sdf=False)
    p = scenes.make_params(default_params, sc.material)
    g = scenes.setup_sim(Sim(p, sc.positions, sc.strand_offsets), sc)
    for s in range(5):
        g.solve_next_step(DT)
    x, v = g.get_pos_vel()
    q = g.get_q()
    assert np.isfinite(x).all() and np.isfinite(v).all() and np.isfinite(q).all()
    roots = sc.roots
    assert np.array_equal(x[roots].view(np.uint32), sc.positions[roots].view(np.uint32))  # pinned roots never move
    np.testing.assert_allclose(np.linalg.norm(q.astype(np.float64), axis=1), 1.0, atol=1e-5)
    # segment lengths stay within 1% of rest (stretching is stiff)
    rest = g.get_rest()["l_rest"]
    tips = sc.strand_offsets[1:] - 1
    mask = np.ones(len(x), bool)
    mask[tips] = False
    seg = np.linalg.norm(x[1:] - x[:-1], axis=1)[mask[:-1]]
    assert np.abs(seg / rest[:-1][mask[:-1]] - 1).max() < 1e-2
    # a strand sample matches the oracle run on that strand alone
    s0, s1 = int(sc.strand_offsets[12345]), int(sc.strand_offsets[12346])
    sub = scenes.Scene("one", sc.positions[s0:s1].copy(), np.array([0, s1 - s0], np.int32))
    o = scenes.setup_sim(OracleSim(p, sub.positions, sub.strand_offsets), sub)
    for s in range(5):
        o.solve_next_step(DT)
    op, ov = o.get_pos_vel()
    parity.check("headline-strand-12345", {"x": x[s0:s1], "v": v[s0:s1], "q": q[s0:s1]},
                 {"x": op, "v": ov, "q": o.get_q()}, {"x": sub.positions})


def test_set_kstretch_and_kbend_match_oracle():
    """Sim::setKStretch / Sim::setKBend (YBSimulator.cu:276-298), the
    UseYoungsModule = false path of the SCR solver: rest data bit for bit,
    then 20 steps."""
    sc = scenes.config_groom(300, 64, with_sdf=False)
    p = scenes.make_params(default_params, sc.material)
    sims = []
    for cls in (Sim, OracleSim):
        s_ = cls(p, sc.positions, sc.strand_offsets)
        s_.configure(sc.material.density)
        s_.set_kstretch(2.5e-5)
        s_.set_kbend(1.5e-9)
        s_.pin(sc.roots)
        s_.upload()
        sims.append(s_)
    g, o = sims
    rg, ro = g.get_rest(), o.get_rest()
    for k in rg:
        assert np.array_equal(np.ascontiguousarray(rg[k]).view(np.uint32), np.ascontiguousarray(ro[k]).view(np.uint32)), k
    q0 = g.get_q()
    for _ in range(20):
        g.solve_next_step(DT)
        o.solve_next_step(DT)
    gp, gv = g.get_pos_vel()
    op, ov = o.get_pos_vel()
    parity.check("set-kstretch-kbend", {"x": gp, "v": gv, "q": g.get_q()}, {"x": op, "v": ov, "q": o.get_q()},
                 {"x": sc.positions, "q": q0})


def test_upload_meta_validation_codes():
    """uploadMeta's throws (YBSimulator.cu:326-332) as status codes; a
    rejected parameter set leaves the handle's parameters unchanged."""
    sc = scenes.config_groom(20, 32, with_sdf=False)
    p = scenes.make_params(default_params, sc.material)
    g = scenes.setup_sim(Sim(p, sc.positions, sc.strand_offsets), sc)
    bad = scenes.make_params(default_params, sc.material)
    bad.detection_scaler = 1.0
    with pytest.raises(KscrError, match="DETECTION_SCALER"):
        g.set_params(bad)
    assert g.get_params().detection_scaler == pytest.approx(p.detection_scaler)
    bad = scenes.make_params(default_params, sc.material)
    bad.barrier_thickness = 5e-3  # 2 * minSegLen < 2 * radius + barrier
    bad.schedule = 1
    with pytest.raises(KscrError, match="SEGMENT_TOO_SHORT"):
        g.set_params(bad)
    kept = g.get_params()
    assert kept.barrier_thickness == pytest.approx(p.barrier_thickness) and kept.schedule == 0
    g.solve_next_step(DT)  # still steps with the kept parameters
    short = scenes.cantilever(8, 4e-4)
    h = Sim(default_params(), short.positions, short.strand_offsets)
    h.configure(1300.0)
    with pytest.raises(KscrError, match="SEGMENT_TOO_SHORT"):
        h.upload()


def test_groom_10k_x_100_sdf128_100_steps():
    """The synthetic groom at its full size: 10k strands x 100 vertices, 128^3
    sphere SDF, 100 steps of 1/600 s, word for word against the FP32
    restatement over the whole groom.  The FP64 error is stated on the first
    1000 strands (strands are independent without contacts)."""
    sc = scenes.config_groom(10_000, 100, sdf_dim=128)
    label = "groom10k-sdf128"
    print(f"[gpu-parity] {label} verts={sc.num_verts} strands={sc.num_strands} steps=100", flush=True)
    g, o, _ = _pair(sc)
    o.set_threads(16)
    sub = scenes.Scene("groom10k-first1000", sc.positions[:100_000].copy(), sc.strand_offsets[:1001].copy(),
                       material=sc.material)
    sub.sdf = sc.sdf
    p = scenes.make_params(default_params, sc.material)
    o64 = scenes.setup_sim(OracleSim(p, sub.positions, sub.strand_offsets, precision=64), sub)
    o64.set_threads(16)
    start = {"x": sc.positions, "q": g.get_q()}
    for s in range(100):
        g.solve_next_step(DT)
        o.solve_next_step(DT)
        o64.solve_next_step(DT)
        if (s + 1) % 20 == 0:
            print(f"[gpu-parity] {label}: step {s + 1}", flush=True)
    gp, gv = g.get_pos_vel()
    op, ov = o.get_pos_vel()
    parity.check(label, {"x": gp, "v": gv, "q": g.get_q()}, {"x": op, "v": ov, "q": o.get_q()}, start)
    p64 = o64.get_pos_vel(f64=True)[0]
    print(f"[gpu-parity] {label} first 1000 strands, FP32 vs FP64 pos: rel {parity.scale_rel(gp[:100_000], p64):.3e}, "
          f"disp-rel {parity.disp_rel(gp[:100_000], p64, sub.positions):.3e}", flush=True)
