"""Statics pins of the oracle's stretch/shear and orientation solves.

The checks here compare the oracle with neither itself nor its adjoint.  They
use closed forms of the discrete Stable Cosserat Rods energy (Hsu, Wang, Wu,
Yuksel, "Stable Cosserat Rods", SIGGRAPH 2025), written here from the paper's
definitions in numpy.

The energy per strand:

    E(x, q) = sum_i (k_i / 2) |(x_{i+1} - x_i) / l_i - R(q_i) e1|^2
            + sum_i (|r_i| - |<conj(q_i) q_{i+1}, r_i>|)
            - sum_i m_i g . x_i

- The first sum is the stretch/shear energy of segment i (stiffness kStretch,
  rest length lRest; YarnBall/sim/cosserat.cu:146-160).
- The second is the bend/twist energy of the joint between frames i and i+1.
  The rest Darboux rotor r_i has norm = bend stiffness (YarnBall.h:55).  The
  absolute value is the rotor sign symmetry q -> -q.
- The last is gravity.

With the roots pinned and the velocity decay set to 1/h, every substep starts
from rest.  The VBD iterations then relax the strand into a static equilibrium
of E.  At that state:

  (1) E is stationary in every free position and every free frame.  This is
      checked by central differences of E in numpy.
  (2) A horizontal clamped cantilever has the tip sag of a discrete linear
      beam: lumped vertex loads m_i g, and joint springs of stiffness |r|/4.
      That stiffness comes from |r| (1 - cos(theta/2)) ~ |r| theta^2 / 8.
  (3) A strand hanging straight down stretches each segment by
      T_i l_i^2 / k_i, where T_i is the weight below the segment.  This
      follows from dE/dx_{i+1} = (k_i / l_i) c_i = T_i.

These catch a wrong weight of the stretch coupling in the orientation solve
(YarnBall/sim/cosserat.cu:310-311).  A factor -k instead of -2k doubles the
effective bend stiffness: it halves the sag in (2) and breaks (1) in the frames.
They also catch a force/Hessian-consistent change of the stretch scale
(kappa = k instead of k/l): (3) fails by the factor l.
"""
import numpy as np
import pytest

from kamisimulator_amd import scenes
from kamisimulator_amd._lib import default_params
from oracle.pyoracle import OracleSim

DT = np.float32(1 / 600)


# ---------------------------------------------------------------- quaternions
def qmul(a, b):
    av, aw = a[..., :3], a[..., 3:4]
    bv, bw = b[..., :3], b[..., 3:4]
    return np.concatenate([aw * bv + bw * av + np.cross(av, bv), aw * bw - np.sum(av * bv, -1, keepdims=True)], -1)


def qconj(a):
    return np.concatenate([-a[..., :3], a[..., 3:]], -1)


def rotate_e1(q):
    e1 = np.zeros(q.shape[:-1] + (4,))
    e1[..., 0] = 1.0
    return qmul(qmul(q, e1), qconj(q))[..., :3]


def qexp(w):
    a = np.linalg.norm(w)
    if a == 0:
        return np.array([0.0, 0.0, 0.0, 1.0])
    return np.concatenate([np.sin(a / 2) * w / a, [np.cos(a / 2)]])


# ---------------------------------------------------------------- the energy
class ScrEnergy:
    """The discrete SCR energy of one strand (module docstring), in float64."""

    def __init__(self, k, l, r, m, g):
        self.k, self.l, self.r, self.m, self.g = k, l, r, m, np.asarray(g, np.float64)

    def __call__(self, x, q):
        n = len(x)
        c = (x[1:] - x[:-1]) / self.l[:n - 1, None] - rotate_e1(q[:n - 1])
        e = 0.5 * np.sum(self.k[:n - 1] * np.sum(c * c, 1))
        d = np.sum(qmul(qconj(q[:n - 2]), q[1:n - 1]) * self.r[:n - 2], 1)
        e += np.sum(np.linalg.norm(self.r[:n - 2], axis=1) - np.abs(d))
        return e - np.sum(self.m * (x @ self.g))

    def gradients(self, x, q, free_x, free_q, hx=1e-7, hq=1e-6):
        """Central differences: dE/dx at free vertices, and dE/dw at free frames
        for the rotation q <- exp(w) q."""
        gx = np.zeros((len(free_x), 3))
        for a, i in enumerate(free_x):
            for c in range(3):
                xp, xm = x.copy(), x.copy()
                xp[i, c] += hx
                xm[i, c] -= hx
                gx[a, c] = (self(xp, q) - self(xm, q)) / (2 * hx)
        gq = np.zeros((len(free_q), 3))
        for a, i in enumerate(free_q):
            for c in range(3):
                w = np.zeros(3)
                w[c] = hq
                qp, qm = q.copy(), q.copy()
                qp[i] = qmul(qexp(w), q[i])
                qm[i] = qmul(qexp(-w), q[i])
                gq[a, c] = (self(x, qp) - self(x, qm)) / (2 * hq)
        return gx, gq


# ---------------------------------------------------------------- relaxation
def relax(sc, young_gpa, max_steps=40_000, chunk=250):
    """Quasi-static FP64 oracle run: roots pinned (and clamped if the scene
    says so), velocity decay 1/h so that each substep starts from rest;
    steps until no position moves by more than 1e-16 m in a chunk of steps."""
    sc.material = scenes.Material(young_gpa=young_gpa, num_itr=10)
    p = scenes.make_params(default_params, sc.material)
    S = int(np.ceil(DT / np.float32(p.max_h)))
    p.velocity_decay = float(np.float32(1.0) / (DT / np.float32(S)))
    o = scenes.setup_sim(OracleSim(p, sc.positions, sc.strand_offsets, precision=64), sc)
    prev = o.get_pos_vel(f64=True)[0]
    moved = np.inf
    for _ in range(max_steps // chunk):
        for _ in range(chunk):
            o.advance(DT)
        x = o.get_pos_vel(f64=True)[0]
        moved = np.abs(x - prev).max()
        prev = x
        if moved < 1e-16:
            break
    rest = o.get_rest()
    im = rest["inv_mass"].astype(np.float64)
    m = np.where(im > 0, 1.0 / np.where(im > 0, im, 1.0), 0.0)
    en = ScrEnergy(rest["k_stretch"].astype(np.float64), rest["l_rest"].astype(np.float64),
                   rest["q_rest"].astype(np.float64), m, p.gravity)
    return prev, o.get_q(f64=True), en, moved


@pytest.fixture(scope="module")
def cantilever_equilibrium():
    n = 16
    sc = scenes.cantilever(n, 1e-3, clamp=True)
    x, q, en, moved = relax(sc, young_gpa=3e-3)
    return n, sc, x, q, en, moved


def test_cantilever_equilibrium_is_stationary(cantilever_equilibrium):
    n, sc, x, q, en, moved = cantilever_equilibrium
    assert moved < 1e-16, f"not relaxed: last chunk moved {moved}"
    free_x = list(range(1, n))      # vertex 0 is pinned
    free_q = list(range(1, n - 1))  # frame 0 is clamped; the tip has no segment
    gx, gq = en.gradients(x, q, free_x, free_q)
    weight = en.m[1] * 9.81
    root_moment = en.m.sum() * 9.81 * (n - 1) * 1e-3
    # the oracle settles where E has no slope, in positions and frames
    ex = np.abs(gx).max() / weight
    eq = np.abs(gq).max() / root_moment
    # the same state with the frames' bend terms doubled (the equilibrium a
    # half-weight stretch coupling in the orientation solve would settle in)
    # has a slope of the order of the bend torques
    en2 = ScrEnergy(en.k, en.l, 2 * en.r, en.m, en.g)
    _, gq2 = en2.gradients(x, q, [], free_q)
    eq2 = np.abs(gq2).max() / root_moment
    print(f"stationarity: |dE/dx|/mg {ex:.2e}, |dE/dq|/M_root {eq:.2e} (doubled bend weight: {eq2:.2e})")
    assert ex < 1e-6
    assert eq < 1e-5
    assert eq2 > 1e-2


def test_cantilever_sag_matches_discrete_beam(cantilever_equilibrium):
    n, sc, x, q, en, _ = cantilever_equilibrium
    l = 1e-3
    g = 9.81
    K = np.linalg.norm(en.r[1]) / 4  # joint stiffness: |r| theta^2 / 8 = K theta^2 / 2
    # linear beam: the joint at vertex j (frames j-1 and j) turns by M_j / K,
    # M_j = sum_{k > j} m_k g (k - j) l, and lifts the tip by theta_j (n-1-j) l
    tip = 0.0
    for j in range(1, n - 1):
        M = sum(en.m[k] * g * (k - j) * l for k in range(j + 1, n))
        tip += M / K * (n - 1 - j) * l
    sag = float(sc.positions[-1, 1]) - x[-1, 1]
    print(f"tip sag {sag:.6e} m, discrete linear beam {tip:.6e} m, ratio {sag / tip:.4f}")
    # 0.6 % of the length: geometric nonlinearity and shear stay below 2 %
    assert abs(sag / tip - 1) < 0.02


def test_hanging_strand_stretch():
    n = 12
    pos = np.zeros((n, 3), np.float32)
    pos[:, 1] = -np.arange(n, dtype=np.float64) * 1e-3
    sc = scenes.Scene("hanging", pos, np.array([0, n], np.int32), clamp_root_orientation=True)
    x, q, en, moved = relax(sc, young_gpa=3e-4)
    assert moved < 1e-16
    seg = np.linalg.norm(x[1:] - x[:-1], axis=1)
    tension = np.array([9.81 * en.m[i + 1:].sum() for i in range(n - 1)])
    expect = en.l[:n - 1] * (1 + tension * en.l[:n - 1] / en.k[:n - 1])
    ext = seg - en.l[:n - 1]
    rel = np.abs(ext / (expect - en.l[:n - 1]) - 1).max()
    print(f"extension of segment 0 {ext[0]:.4e} m (closed form {expect[0] - en.l[0]:.4e}); max rel dev {rel:.2e}")
    assert rel < 1e-4
    # straight down, frames along the segments (up to the float32 rounding of
    # the rest rotors, ~1e-7 of a segment per joint)
    np.testing.assert_allclose(x[:, [0, 2]], 0.0, atol=1e-5 * 1e-3)
    np.testing.assert_allclose(rotate_e1(q[:n - 1]), np.tile([0.0, -1.0, 0.0], (n - 1, 1)), atol=1e-5)
