"""Pins of the oracle's SagFree restatement (scr_oracle.hpp: sag_free,
following YB/YBSimulator.cu:480-676), which no reference test or fixture
covers.

* Pass 1 (rest lengths): each new rest length l of a segment with current
  length L and stretch stiffness k balances the weight M g_axis it carries,
  k (L - l) / l^2 = M g_axis, on the root nearer L of that quadratic.  M is the
  reference's running mass (the tip, every vertex between, half the segment's
  lower vertex); the carried mass at the root is the strand's mass within half
  a vertex.
* Pass 2 (frames): on a horizontal strand the stretch pass carries no load,
  the load beyond vertex v is M_v g, and the frame of segment v turns in the
  vertical plane by theta_v = 4 asin(T_v / 2) with
  T_v = M_v g cos(alpha_v) l_v / k_bend_v, alpha_v the tilt accumulated by the
  segments before it (cos(theta / 2) = 1 - T^2 / 2 of the reference's
  closed form); the turns accumulate along the strand.
* Invariants: positions, masses and bend stiffnesses |qRest| unchanged; every
  rest rotor the normalized relative rotor of the new frames; k iterations
  equal k single iterations.
"""
import numpy as np
import pytest

from kamisimulator_amd import scenes
from kamisimulator_amd._lib import default_params
from oracle.pyoracle import OracleSim

F32 = np.float32


def _strand(direction, n=24, l=1e-3):
    d = np.array(direction, float)
    d /= np.linalg.norm(d)
    return (np.arange(n)[:, None] * l * d[None, :]).astype(F32)


def _sim(pos, young_gpa, sag=0, gravity=(0.0, -9.8, 0.0), upload=True):
    m = scenes.Material(young_gpa=young_gpa)
    p = scenes.make_params(default_params, m)
    for a in range(3):
        p.gravity[a] = gravity[a]
    o = OracleSim(p, pos, [0, len(pos)], precision=64)
    o.configure(m.density)
    o.set_young(m.young_gpa, m.deviation)
    o.pin([0])
    if sag:
        o.sag_free([0], sag)
    if upload:
        o.upload()
    return o


def _rot(q, v):
    """Rotate v by unit quaternions q (x, y, z, w), rows."""
    qv, w = q[:, :3], q[:, 3:]
    t = 2 * np.cross(qv, v)
    return v + w * t + np.cross(qv, t)


def test_rest_lengths_balance_the_carried_weight(oracle_lib):
    pos = _strand((0.4, -1.0, 0.3))
    before = _sim(pos, 1e-4).get_rest()
    after = _sim(pos, 1e-4, sag=1).get_rest()
    g = np.array([0.0, -9.8, 0.0], F32).astype(np.float64)
    n = len(pos)
    with np.errstate(divide="ignore"):
        m = 1.0 / before["inv_mass"].astype(np.float64)  # the pinned root (inf) is not used
    axis = pos[1] - pos[0]
    axis = axis.astype(np.float64) / np.linalg.norm(axis.astype(np.float64))
    ga = axis @ g
    L = before["l_rest"].astype(np.float64)
    k = before["k_stretch"].astype(np.float64)
    l_new = after["l_rest"].astype(np.float64)
    carried = m[n - 1]
    for seg in range(n - 2, 0, -1):
        carried += 0.5 * m[seg]
        lhs = k[seg] * (L[seg] - l_new[seg]) / l_new[seg] ** 2
        # float rounding of the new length bounds the residual
        tol = k[seg] * L[seg] / l_new[seg] ** 2 * 2 ** -23 * 4
        assert lhs == pytest.approx(carried * ga, abs=tol), seg
        assert l_new[seg] < L[seg]  # hanging: the rest length shortens
        carried += 0.5 * m[seg]
    # the mass carried by the root segment is the strand's free mass
    assert carried + 0.5 * m[1] == pytest.approx(m[1:].sum() + 0.5 * m[1], rel=1e-6)
    # the change is far above float resolution at this stiffness
    assert (L[1:-1] - l_new[1:-1]).max() > 1e3 * 2 ** -23 * L[1:-1].max()


def test_frames_turn_by_the_weight_torque(oracle_lib):
    pos = _strand((1.0, 0.0, 0.0))
    o0 = _sim(pos, 1e-3)
    o1 = _sim(pos, 1e-3, sag=1)
    r0, r1 = o0.get_rest(), o1.get_rest()
    q0, q1 = o0.get_q(f64=True), o1.get_q(f64=True)
    n = len(pos)
    # the stretch pass carries no load on a horizontal strand
    np.testing.assert_array_equal(r1["l_rest"], r0["l_rest"])
    with np.errstate(divide="ignore"):
        m = 1.0 / r0["inv_mass"].astype(np.float64)  # the pinned root (inf) is not used
    g = 9.8
    l = np.linalg.norm(np.diff(pos.astype(np.float64), axis=0), axis=1)
    k_bend = np.linalg.norm(r0["q_rest"].astype(np.float64), axis=1)
    beyond = m[1:].sum()
    alpha = 0.0
    want = np.zeros(n)
    for v in range(1, n - 1):
        beyond -= 0.5 * m[v]
        T = beyond * g * np.cos(alpha) * l[v] / k_bend[v]
        alpha += 4 * np.arcsin(T / 2)
        want[v] = alpha
        beyond -= 0.5 * m[v]
    e1 = np.tile([1.0, 0.0, 0.0], (n, 1))
    d0, d1 = _rot(q0, e1), _rot(q1, e1)
    tilt = np.arctan2(-d1[:, 1], d1[:, 0]) - np.arctan2(-d0[:, 1], d0[:, 0])
    # the turns point towards gravity (the reference's sign), stay in the
    # vertical plane and accumulate along the strand
    np.testing.assert_allclose(tilt[1:n - 1], want[1:n - 1], rtol=1e-5, atol=1e-9)
    assert np.abs(d1[:, 2]).max() < 1e-12
    assert want[n - 2] > 1e-3


def test_invariants(oracle_lib):
    pos = _strand((0.4, -1.0, 0.3))
    o0 = _sim(pos, 1e-3)
    o1 = _sim(pos, 1e-3, sag=1)
    r0, r1 = o0.get_rest(), o1.get_rest()
    np.testing.assert_array_equal(r1["inv_mass"], r0["inv_mass"])
    np.testing.assert_array_equal(r1["k_stretch"], r0["k_stretch"])
    np.testing.assert_allclose(np.linalg.norm(r1["q_rest"], axis=1), np.linalg.norm(r0["q_rest"], axis=1), rtol=1e-6)
    np.testing.assert_array_equal(o1.get_pos_vel()[0], pos)
    # every rest rotor is the relative rotor of the new frames, scaled by its stiffness
    q = o1.get_q(f64=True)
    qr = r1["q_rest"].astype(np.float64)
    for i in range(1, len(pos) - 2):
        a, b = q[i], q[i + 1]
        rel = np.array([a[3] * b[0] - b[3] * a[0] - (a[1] * b[2] - a[2] * b[1]),
                        a[3] * b[1] - b[3] * a[1] - (a[2] * b[0] - a[0] * b[2]),
                        a[3] * b[2] - b[3] * a[2] - (a[0] * b[1] - a[1] * b[0]),
                        a[3] * b[3] + a[:3] @ b[:3]])  # conj(a) (x) b
        rel /= np.linalg.norm(rel)
        np.testing.assert_allclose(qr[i] / np.linalg.norm(qr[i]), rel, atol=2e-6)
    # k iterations = k single iterations
    o2 = _sim(pos, 1e-3, sag=2)
    o11 = _sim(pos, 1e-3, sag=1, upload=False)
    o11.sag_free([0], 1)
    o11.upload()
    r2, r11 = o2.get_rest(), o11.get_rest()
    for key in ("l_rest", "q_rest"):
        np.testing.assert_array_equal(r2[key], r11[key])
    np.testing.assert_array_equal(o2.get_q(), o11.get_q())
