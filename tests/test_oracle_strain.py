"""Pins the oracle's strain limiting (scr_oracle.hpp apply_strain_limit,
following SDF/cudaSDFCollision.cu:176-199) by closed forms.

The step is made motionless (no iterations, g = 0, v = 0: every substep moves
x by h v = 0) and the body SDF lies far away (phi > 0 everywhere), so the
strain pass is the only thing that moves a vertex:

* a straight strand stretched uniformly along its line comes out with
  vertex k at root + (l0_0 + ... + l0_{k-1}) d: the walk resets each segment
  to its rest length along the segment;
* on a compressed line the walk is scalar and folds back once the shortfall
  exceeds a rest length (each vertex goes to the side of the moved
  predecessor where its old position lies);
* a strand whose segments are within the tolerance is left bit for bit;
* a segment exactly at the tolerance is left (the test is strict, l - l0 >
  err l0);
* roots and velocities never move;
* with the SDF pass off (advance, or solve_next_step without a body) the
  strain pass does not run;
* rest lengths default to the segment lengths when the limit is set (the
  SDF solver's InitAdditionalHairData), so an unmoved groom is untouched.
"""
import numpy as np
import pytest

from kamisimulator_amd import scenes
from kamisimulator_amd._lib import default_params
from oracle.pyoracle import OracleSim

F32 = np.float32
DT = F32(1 / 600)


def _params():
    p = scenes.make_params(default_params, scenes.Material())
    p.num_itr = 0
    for a in range(3):
        p.gravity[a] = 0.0
    p.velocity_decay = 0.0
    return p


def _far_sdf():
    # phi = 10 + z > 0 on the whole grid: the SDF pass never projects
    nrm = np.array([0.0, 0.0, 1.0])
    return scenes.sdf_grid(lambda P: (P @ nrm + 10.0, np.broadcast_to(nrm, P.shape)), (-1, -1, -1), (1, 1, 1),
                           (8, 8, 8))


def _sim(pos, off, precision, rest=None, err=0.1, sdf=True):
    o = OracleSim(_params(), pos, off, precision=precision)
    o.configure(1300.0)
    o.set_young(3.0, 1.0)
    o.pin(np.asarray(off[:-1], np.int32))
    o.upload()
    o.set_pos_vel(pos, np.zeros_like(pos))
    if sdf:
        o.set_sdf(_far_sdf(), threshold=0.0)
    o.set_strain_limit(True, err, rest)
    return o


def _line(root, d, lengths):
    d = np.asarray(d, np.float64)
    d /= np.linalg.norm(d)
    pts = [np.asarray(root, np.float64)]
    for l in lengths:
        pts.append(pts[-1] + l * d)
    return np.array(pts), d


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("factor", [1.25, 2.0, 1.11])
def test_straight_strand_reset_to_rest_lengths(oracle_lib, precision, factor):
    rng = np.random.default_rng(3)
    nv = 40
    l0 = rng.uniform(0.8e-3, 1.2e-3, nv - 1)
    root = np.array([0.01, -0.02, 0.03])
    d = np.array([0.3, -0.5, 0.8])
    pts, dn = _line(root, d, factor * l0)
    pos = pts.astype(F32)
    rest = np.zeros(nv, F32)
    rest[:-1] = l0.astype(F32)
    o = _sim(pos, [0, nv], precision, rest=rest, err=0.1)
    o.solve_next_step(DT)
    P, V = o.get_pos_vel(f64=True)
    # closed form: vertex k at root + (sum of the first k rest lengths) d
    want = root + np.concatenate([[0.0], np.cumsum(rest[:-1].astype(np.float64))])[:, None] * dn
    tol = 5e-9 if precision == 64 else 2e-6  # FP32 positions: ~1e-7 relative per step of the walk
    np.testing.assert_allclose(P, want, rtol=0, atol=tol)
    assert np.array_equal(P[0], pos[0].astype(np.float64))  # the root stays
    assert not V.any()  # velocities untouched


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("factor,axis", [(0.75, 0), (0.3, 1), (0.89, 2)])
def test_compressed_line_walk(oracle_lib, precision, factor, axis):
    """A uniformly compressed straight strand: each vertex is put at its rest
    length from the already moved one, on the side where its old position
    lies.  Once the accumulated shortfall exceeds a rest length, an old vertex
    lies behind the moved one and the segment folds back towards the root.
    On the line the walk is scalar: t_{k+1} = t_k + l0_k sign(s_{k+1} - t_k)
    when ||s_{k+1} - t_k| - l0_k| > err l0_k, else s_{k+1}.  The line runs
    along a coordinate axis, so the FP32 inputs lie on it exactly (off an
    oblique line their rounding tilts a fold's direction when an old vertex
    lies micrometres from the moved one)."""
    rng = np.random.default_rng(5)
    nv = 40
    l0 = rng.uniform(0.8e-3, 1.2e-3, nv - 1)
    root = np.array([0.01, -0.02, 0.03])
    pts, dn = _line(root, np.eye(3)[axis], factor * l0)
    pos = pts.astype(F32)
    rest = np.zeros(nv, F32)
    rest[:-1] = l0.astype(F32)
    o = _sim(pos, [0, nv], precision, rest=rest, err=0.1)
    o.solve_next_step(DT)
    P, V = o.get_pos_vel(f64=True)
    p64 = pos.astype(np.float64)
    s_old = (p64 - p64[0]) @ dn  # the FP32 inputs' positions along the line
    r = rest[:-1].astype(np.float64)
    t = [0.0]
    for k in range(nv - 1):
        gap = s_old[k + 1] - t[k]
        t.append(t[k] + r[k] * np.sign(gap) if abs(abs(gap) - r[k]) > 0.1 * r[k] else s_old[k + 1])
    want = p64[0] + np.array(t)[:, None] * dn
    if factor < 0.8:
        assert np.any(np.diff(t) < 0)  # the walk folds back somewhere
    tol = 5e-9 if precision == 64 else 2e-6
    np.testing.assert_allclose(P, want, rtol=0, atol=tol)
    assert not V.any()


@pytest.mark.parametrize("precision", [32, 64])
def test_strand_within_tolerance_is_untouched(oracle_lib, precision):
    rng = np.random.default_rng(4)
    nv = 30
    l0 = rng.uniform(0.8e-3, 1.2e-3, nv - 1)
    # every segment within +-9 % of its rest length (tolerance 10 %), on a curve
    f = rng.uniform(0.91, 1.09, nv - 1)
    dirs = rng.normal(size=(nv - 1, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    pts = np.concatenate([[np.zeros(3)], np.cumsum((f * l0)[:, None] * dirs, axis=0)])
    pos = pts.astype(F32)
    rest = np.zeros(nv, F32)
    rest[:-1] = l0.astype(F32)
    o = _sim(pos, [0, nv], precision, rest=rest, err=0.1)
    o.solve_next_step(DT)
    P, _ = o.get_pos_vel()
    assert np.array_equal(P.view(np.uint32), pos.view(np.uint32))


def test_only_violating_segments_move(oracle_lib):
    """One segment of a straight strand 50 % too long: the vertices below it
    stay; its upper vertex moves back to the rest length, which stretches the
    next segment by the same 0.5 mm, and so on up the strand, so vertex k ends
    at k mm.  With a tolerance above 50 % nothing moves."""
    nv = 10
    l0 = np.full(nv - 1, 1e-3)
    lens = l0.copy()
    lens[4] = 1.5e-3
    pts, d = _line(np.zeros(3), [1.0, 0.0, 0.0], lens)
    pos = pts.astype(F32)
    rest = np.zeros(nv, F32)
    rest[:-1] = l0.astype(F32)
    o = _sim(pos, [0, nv], 64, rest=rest, err=0.1)
    o.solve_next_step(DT)
    P, _ = o.get_pos_vel(f64=True)
    np.testing.assert_allclose(P[:, 0], np.arange(nv) * 1e-3, rtol=0, atol=1e-9)  # FP32 inputs
    assert not P[:, 1:].any()
    # with a tolerance above the 50 % strain nothing moves
    o = _sim(pos, [0, nv], 64, rest=rest, err=0.6)
    o.solve_next_step(DT)
    P, _ = o.get_pos_vel()
    assert np.array_equal(P.view(np.uint32), pos.view(np.uint32))


def test_strict_threshold(oracle_lib):
    """|l - l0| == err l0 is not a violation (the reference tests >)."""
    nv = 3
    pos = np.array([[0, 0, 0], [1.5, 0, 0], [3.0, 0, 0]], F32)
    rest = np.array([1.0, 1.0, 0.0], F32)
    o = _sim(pos, [0, nv], 32, rest=rest, err=0.5)  # |1.5 - 1| = 0.5 * 1: kept
    o.solve_next_step(DT)
    P, _ = o.get_pos_vel()
    assert np.array_equal(P.view(np.uint32), pos.view(np.uint32))


@pytest.mark.parametrize("precision", [32, 64])
def test_several_strands_and_default_rest_lengths(oracle_lib, precision):
    sc = scenes.config_groom(40, 24, with_sdf=False)
    pos = sc.positions
    # default rest lengths = the current segment lengths: an unmoved groom
    # stays (the tolerance only has to cover the last-bit difference between
    # the host's unfused length and the walk's FMA dot, as in the reference)
    o = _sim(pos, sc.strand_offsets, precision, rest=None, err=1e-6)
    o.solve_next_step(DT)
    P, _ = o.get_pos_vel()
    assert np.array_equal(P.view(np.uint32), pos.view(np.uint32))
    # the last segment of every curved strand stretched by 1.3 along itself:
    # only the tips move, back onto their initial positions
    off = sc.strand_offsets
    stretched = pos.astype(np.float64).copy()
    tips = off[1:] - 1
    stretched[tips] = stretched[tips - 1] + 1.3 * (stretched[tips] - stretched[tips - 1])
    o = _sim(stretched.astype(F32), off, precision, rest=None, err=0.1)  # rest = stretched lengths
    o.set_strain_limit(True, 0.1, _seg_lengths(pos, off))                  # rest = initial lengths
    o.solve_next_step(DT)
    P, _ = o.get_pos_vel(f64=True)
    keep = np.ones(len(pos), bool)
    keep[tips] = False
    assert np.array_equal(P[keep], stretched.astype(F32).astype(np.float64)[keep])
    np.testing.assert_allclose(P[tips], pos[tips].astype(np.float64), rtol=0, atol=2e-8)


def _seg_lengths(pos, off):
    rest = np.zeros(len(pos), F32)
    for s in range(len(off) - 1):
        a, b = off[s], off[s + 1]
        seg = pos[a + 1:b].astype(np.float64) - pos[a:b - 1].astype(np.float64)
        rest[a:b - 1] = np.linalg.norm(seg, axis=1).astype(F32)
    return rest


def test_not_run_without_the_sdf_pass(oracle_lib):
    nv = 12
    pts, _ = _line(np.zeros(3), [0.0, 1.0, 0.0], np.full(nv - 1, 2e-3))
    pos = pts.astype(F32)
    rest = np.full(nv, 1e-3, F32)
    for sdf, step in ((True, "advance"), (False, "solve_next_step")):
        o = _sim(pos, [0, nv], 32, rest=rest, err=0.1, sdf=sdf)
        getattr(o, step)(DT)
        P, _ = o.get_pos_vel()
        assert np.array_equal(P.view(np.uint32), pos.view(np.uint32)), step
    # switched off
    o = _sim(pos, [0, nv], 32, rest=rest, err=0.1)
    o.set_strain_limit(False)
    o.solve_next_step(DT)
    P, _ = o.get_pos_vel()
    assert np.array_equal(P.view(np.uint32), pos.view(np.uint32))
