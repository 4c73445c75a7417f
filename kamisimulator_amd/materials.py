"""Per-strand materials (Young's modulus and rest-curvature scale) and their
derivatives, for the differentiable step.

Per vertex i of strand s (reference setup: Sim::setYoung,
src/solvers/StableCosseratRods/YarnBall/YBSimulator.cu:300-320):
    kStretch_i = E_s * pi * mmrad^2                         (mmrad = radius * 1000)
    qRest_i    = bt_i * normalize(kappa_s * r_i)
    bt_i       = kStretch_i * deviation * mmrad^2 / (4 l_i^2)
where r_i is the unit rest rotor q_i^* q_{i+1} of the initial shape and
kappa * r is the reference's rotor power (angle scaling,
KE/includes/modules/Rotor.h:221-228).  Everything here is float64 host math.
"""
import numpy as np


def rotor_power(kappa, r):
    """kappa * r for unit rotors r[..., 4] (x, y, z, w); returns (value, d/dkappa)."""
    r = np.asarray(r, np.float64)
    kappa = np.asarray(kappa, np.float64)
    qv = r[..., :3]
    w = np.clip(r[..., 3], -1.0, 1.0)
    qn = np.linalg.norm(qv, axis=-1)
    alpha = np.arccos(w)
    na = kappa * alpha
    safe = np.where(qn > 0, qn, 1.0)
    axis = qv / safe[..., None]
    val = np.concatenate([axis * np.sin(na)[..., None], np.cos(na)[..., None]], axis=-1)
    dval = np.concatenate([axis * (alpha * np.cos(na))[..., None], (-alpha * np.sin(na))[..., None]], axis=-1)
    ident = (w == 1.0) | (qn == 0)
    val[ident] = [0, 0, 0, 1]
    dval[ident] = 0
    return val, dval


def strand_materials(rest_dir, l_rest, strand_offsets, young_gpa, kappa, radius, deviation=1.0):
    """Per-vertex (kStretch[n], qRest[n, 4]) and their derivatives with respect
    to each vertex's strand E (dks_dE[n], dqr_dE[n, 4]) and kappa (dqr_dk[n, 4]).

    rest_dir: unit rest rotors r_i (n, 4); l_rest: rest lengths (n);
    young_gpa, kappa: per strand (S)."""
    off = np.asarray(strand_offsets)
    counts = np.diff(off)
    E = np.repeat(np.asarray(young_gpa, np.float64), counts)
    K = np.repeat(np.asarray(kappa, np.float64), counts)
    mmrad = float(radius) * 1000.0
    c1 = np.pi * mmrad * mmrad
    ks = E * c1
    l = np.asarray(l_rest, np.float64)
    bt = ks * deviation * mmrad * mmrad * 0.25 / (l * l)
    val, dval = rotor_power(K, rest_dir)
    qr = bt[:, None] * val
    return dict(k_stretch=ks, q_rest=qr, dks_dE=np.full_like(ks, c1), dqr_dE=(bt / E)[:, None] * val,
                dqr_dk=bt[:, None] * dval)


def strand_gradient(mat, strand_offsets, ksb, qrb):
    """Chain rule from per-vertex adjoints to per-strand dL/dE and dL/dkappa."""
    off = np.asarray(strand_offsets)
    gE_v = ksb * mat["dks_dE"] + np.sum(qrb * mat["dqr_dE"], axis=1)
    gk_v = np.sum(qrb * mat["dqr_dk"], axis=1)
    return np.add.reduceat(gE_v, off[:-1]), np.add.reduceat(gk_v, off[:-1])


def rest_directions(q_rest):
    """Unit rest rotors from the rest rotors a setup produced."""
    q = np.asarray(q_rest, np.float64)
    n = np.linalg.norm(q, axis=1, keepdims=True)
    return np.where(n > 0, q / np.where(n > 0, n, 1), np.array([0, 0, 0, 1.0]))
