"""The glue between connected vertices (Vertex::connectionIndex,
YarnBall.h:56) in the oracle, pinned independently of its formulas.

The reference glues a closed curve's two ends (createFromCurves,
YarnBall/io/reader.cpp:40-44) and nearby curve endpoints (glueEndpoints,
YarnBall/io/jsonBuilder.cpp:163-188).  Its position solve adds a spring of
stiffness 40 between a vertex and its connected vertex
(YarnBall/sim/cosserat.cu:123-130).  The narrow phase skips segment pairs
that the glue joins (YarnBall/sim/collision.cu:44-49).

Checks:
- the glue is the energy 20 |x_i - x_j|^2 per glued pair: a clamped
  cantilever whose tip is glued to a pinned anchor settles at a stationary
  point of the discrete SCR energy plus that term (numpy central
  differences);
- one substep of a vertex glued to a pinned vertex equals a numpy restatement
  of the per-vertex 3x3 Newton step, damping included (the inertia Hessian
  fills all six packed entries, KE/includes/modules/SymMat.h:16-18);
- a closed ring whose glued ends coincide stays at rest without gravity;
  under gravity, with one vertex pinned, its ends stay within the gap the
  spring allows;
- contact lists skip glued segment pairs.
"""
import numpy as np
import pytest

from kamisimulator_amd import scenes
from kamisimulator_amd._lib import default_params
from oracle.pyoracle import OracleSim
from tests.test_oracle_statics import ScrEnergy, qexp, qmul

F32 = np.float32
DT = F32(1 / 600)


def _quasi_static(o, dt=DT, max_steps=40_000, chunk=250):
    prev = o.get_pos_vel(f64=True)[0]
    moved = np.inf
    for _ in range(max_steps // chunk):
        for _ in range(chunk):
            o.advance(dt)
        x = o.get_pos_vel(f64=True)[0]
        moved = np.abs(x - prev).max()
        prev = x
        if moved < 1e-16:
            break
    return prev, moved


def test_glued_cantilever_equilibrium_is_stationary(oracle_lib):
    n = 12
    l = 1e-3
    pos = np.zeros((n + 2, 3), np.float32)
    pos[:n, 0] = np.arange(n) * l
    # anchor strand: two pinned vertices, the first 0.3 mm above the tip
    pos[n] = [(n - 1) * l, 3e-4, 0.0]
    pos[n + 1] = [(n - 1) * l, 3e-4, l]
    off = np.array([0, n, n + 2], np.int32)
    m = scenes.Material(young_gpa=3e-3, num_itr=10)
    p = scenes.make_params(default_params, m)
    S = int(np.ceil(DT / np.float32(p.max_h)))
    p.velocity_decay = float(np.float32(1.0) / (DT / np.float32(S)))
    o = OracleSim(p, pos, off, precision=64)
    o.configure(m.density)
    o.set_young(m.young_gpa, m.deviation)
    o.pin([0, n, n + 1])
    o.fix_orientation([0])
    o.upload()
    conn = np.full(n + 2, -1, np.int32)
    conn[n - 1], conn[n] = n, n - 1
    o.set_connections(conn)
    x, moved = _quasi_static(o)
    assert moved < 1e-16
    q = o.get_q(f64=True)
    rest = o.get_rest()
    im = rest["inv_mass"].astype(np.float64)
    mass = np.where(im > 0, 1.0 / np.where(im > 0, im, 1.0), 0.0)
    en = ScrEnergy(rest["k_stretch"][:n].astype(np.float64), rest["l_rest"][:n].astype(np.float64),
                   rest["q_rest"][:n].astype(np.float64), mass[:n], p.gravity)

    def total(xa, qa):
        d = xa[n - 1] - x[n]
        return en(xa, qa) + 20.0 * d @ d

    gx = np.zeros((n - 1, 3))
    for a, i in enumerate(range(1, n)):
        for c in range(3):
            xp, xm = x[:n].copy(), x[:n].copy()
            xp[i, c] += 1e-7
            xm[i, c] -= 1e-7
            gx[a, c] = (total(xp, q[:n]) - total(xm, q[:n])) / 2e-7
    gq = np.zeros((n - 2, 3))
    for a, i in enumerate(range(1, n - 1)):
        for c in range(3):
            w = np.zeros(3)
            w[c] = 1e-6
            qp, qm = q[:n].copy(), q[:n].copy()
            qp[i] = qmul(qexp(w), q[i])
            qm[i] = qmul(qexp(-w), q[i])
            gq[a, c] = (total(x[:n], qp) - total(x[:n], qm)) / 2e-6
    pull = 40.0 * np.linalg.norm(x[n - 1] - x[n])
    print(f"glue force {pull:.3e} N, |dE/dx| {np.abs(gx).max():.3e}, |dE/dq| {np.abs(gq).max():.3e}, "
          f"tip lift {x[n - 1, 1] - pos[n - 1, 1]:.3e} m")
    # the spring carries a force of the order of the strand's weight and lifts the tip
    assert pull > 0.1 * mass[:n].sum() * 9.81
    assert x[n - 1, 1] > pos[n - 1, 1]
    assert np.abs(gx).max() < 1e-6 * pull
    assert np.abs(gq).max() < 1e-6 * pull * (n - 1) * l


def test_one_substep_closed_form(oracle_lib):
    """Vertex 1 (tip of strand 0-1, root pinned and clamped) glued to pinned
    vertex 3: one substep of 1/3000 s, numItr iterations, in numpy."""
    l = 2.0 ** -10
    pos = np.array([[0, 0, 0], [l, 0, 0], [l, 2.0 ** -12, 0], [l, 2.0 ** -12, l]], F32)
    g = np.array([0.0, -9.81, 0.0])
    m = scenes.Material(num_itr=4, gravity=tuple(g))
    p = scenes.make_params(default_params, m)
    p.damping = 1e-3
    o = OracleSim(p, pos, [0, 2, 4], precision=64)
    o.configure(m.density)
    o.set_young(m.young_gpa, m.deviation)
    o.pin([0, 2, 3])
    o.fix_orientation([0])
    o.upload()
    o.set_connections([-1, 3, -1, 1])
    rest = o.get_rest()
    dt = F32(1 / 3000)
    o.advance(dt)  # one substep (dt <= max_h)
    P, V = o.get_pos_vel(f64=True)

    h = float(F32(dt))
    inv_m = float(rest["inv_mass"][1])
    kst, lr = float(rest["k_stretch"][0]), float(rest["l_rest"][0])
    zeta = float(F32(p.damping)) / h
    a = 1.0 / (inv_m * h * h)
    dseg = (kst / lr) / lr
    gext = g.astype(F32).astype(np.float64)
    # from rest: v = u = 0, so the adaptive guess's acceleration is 0 (sigma = 0)
    y = h * h * gext
    dx = np.zeros(3)
    x1, x3 = pos[1].astype(np.float64), pos[3].astype(np.float64)
    H = a * np.ones((3, 3)) + (40.0 * (1 + zeta)) * np.eye(3) + dseg * (1 + zeta) * np.eye(3)
    for _ in range(p.num_itr):
        # segment (0, 1) at rest: c = dx_1 / l (frame 0 is clamped along +x)
        c = dx / lr
        f = a * (y - dx) - 40.0 * ((x1 - x3) + dx + zeta * dx) + (-(kst / lr) * c - zeta * dseg * dx)
        dx = dx + np.linalg.solve(H, f)
    want = x1 + dx
    np.testing.assert_allclose(P[1], want, rtol=0, atol=1e-9 * l)
    # the spring pulls vertex 1 towards vertex 3, against gravity
    assert P[1][1] > pos[1][1]


def _ring(n=40, radius=0.01, y0=0.0):
    th = np.linspace(0.0, 2 * np.pi, n)  # last point = first point
    pos = np.stack([radius * np.cos(th), y0 + radius * np.sin(th), np.zeros(n)], 1)
    pos[-1] = pos[0]
    return pos.astype(F32)


def _ring_sim(pos, gravity, prec=64, pin=None):
    m = scenes.Material(gravity=gravity)
    p = scenes.make_params(default_params, m)
    n = len(pos)
    o = OracleSim(p, pos, [0, n], precision=prec)
    o.configure(m.density)
    o.set_young(m.young_gpa, m.deviation)
    if pin is not None:
        o.pin(pin)
    o.upload()
    conn = np.full(n, -1, np.int32)
    conn[0], conn[n - 1] = n - 1, 0
    o.set_connections(conn)
    return o


def test_closed_ring_rest_invariance(oracle_lib):
    pos = _ring()
    o = _ring_sim(pos, (0.0, 0.0, 0.0))
    q0 = o.get_q(f64=True)
    for _ in range(10):
        o.advance(DT)
    P, _ = o.get_pos_vel(f64=True)
    # FP32 rest data leave residuals of ~1e-7 in the constraints; a sign or
    # scale error in the glue would move the joined ends by O(gap)
    assert np.abs(P - pos.astype(np.float64)).max() < 1e-8
    assert np.abs(o.get_q(f64=True) - q0).max() < 1e-5


def test_ring_stays_closed_under_gravity(oracle_lib):
    pos = _ring()
    n = len(pos)
    top = int(np.argmax(pos[:, 1]))
    o = _ring_sim(pos, (0.0, -9.81, 0.0), pin=[top])
    for _ in range(100):
        o.advance(DT)
    P, _ = o.get_pos_vel(f64=True)
    rest = o.get_rest()
    im = rest["inv_mass"].astype(np.float64)
    weight = np.sum(np.where(im > 0, 1.0 / np.where(im > 0, im, 1.0), 0.0)) * 9.81
    gap = np.linalg.norm(P[0] - P[n - 1])
    print(f"ring gap after 100 steps {gap:.3e} m (spring bound {weight / 40:.3e} m)")
    assert np.isfinite(P).all()
    assert gap <= weight / 40.0
    # the same ring without the glue opens at the joint
    m = scenes.Material(gravity=(0.0, -9.81, 0.0))
    p = scenes.make_params(default_params, m)
    o2 = OracleSim(p, pos, [0, n], precision=64)
    o2.configure(m.density)
    o2.set_young(m.young_gpa, m.deviation)
    o2.pin([top])
    o2.upload()
    for _ in range(100):
        o2.advance(DT)
    P2, _ = o2.get_pos_vel(f64=True)
    assert np.linalg.norm(P2[0] - P2[n - 1]) > 10 * gap


def test_contact_lists_skip_glued_segments(oracle_lib):
    """Two crossing strands glued at their crossing: the narrow phase
    exempts the pairs (a, b) where a glued vertex of segment a is b or b + 1
    (collision.cu:44-49), and lists every other close pair."""
    l = 1e-3
    k = np.arange(8)
    s0 = np.stack([k * l, np.zeros(8), np.zeros(8)], 1)
    s1 = np.stack([3.5 * l * np.ones(8), (k - 3.5) * l, 4e-5 * np.ones(8)], 1)
    pos = np.concatenate([s0, s1]).astype(F32)
    m = scenes.Material()
    p = scenes.contact_params(default_params, m)
    res = []
    for glued in (False, True):
        o = OracleSim(p, pos, [0, 8, 16], precision=32)
        o.configure(m.density)
        o.set_young(m.young_gpa, m.deviation)
        o.pin([0, 8])
        o.upload()
        if glued:
            conn = np.full(16, -1, np.int32)
            conn[3], conn[12] = 12, 3  # vertex 3 of strand 0 and vertex 4 of strand 1
            o.set_connections(conn)
        o.detect_only(DT)
        res.append(o.contacts())
    (n0, c0), (n1, c1) = res

    def pairs(num, cols):
        return {(a, int(b)) for a in range(16) for b in cols[a, :num[a]]}

    plain, glued = pairs(n0, c0), pairs(n1, c1)
    conn = {3: 12, 12: 3}

    def exempt(a, b):
        c = [conn.get(a, -1), conn.get(a + 1, -1)]
        return b in c or b + 1 in c

    assert (3, 11) in plain and (11, 3) in plain  # the crossing is a contact without the glue
    assert glued == {(a, b) for (a, b) in plain if not (exempt(a, b) or exempt(b, a))}
    assert len(glued) < len(plain)
