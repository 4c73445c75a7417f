"""Accuracy of per-strand gradients of the differentiable step, measured
against the oracle.

TEST INFRASTRUCTURE ONLY: used by tests/ and by the CPU leg of bench.py as
the checker of a gradient the HIP library computed.

Strands are independent without contacts.  A gradient computed on the device
for a whole groom can therefore be checked strand by strand: each check
reruns a few strands of the groom alone on the CPU.  The measures are:

- the FP32 oracle gradient of the same rollout.  The device's per-vertex
  adjoints equal it bit for bit.
- the FP64 oracle gradient: the same reverse sweep, in double precision.
- FP64 central finite differences of the loss (h = 1e-6 relative) for E and
  kappa of chosen strands.  These pin the FP64 gradient.
- the branch of the adaptive initial guess's clamp (YarnBall/sim/iteration.cu:
  50) per vertex and substep, in the FP32 and the FP64 rollout.  A branch
  that differs between them makes the FP32 gradient the derivative of a
  different piecewise-smooth map.
"""
import numpy as np

from kamisimulator_amd import materials, scenes

from .pyoracle import OracleSim


def _sim(sc, params, young, kappa, prec, threads):
    o = OracleSim(params, sc.positions, sc.strand_offsets, precision=prec)
    o.set_threads(threads)
    scenes.setup_sim(o, sc)
    rest = o.get_rest()
    mat = materials.strand_materials(materials.rest_directions(rest["q_rest"]), rest["l_rest"], sc.strand_offsets,
                                     young, kappa, params.radius, sc.material.deviation)
    if prec == 64:
        o.set_material64(mat["k_stretch"], mat["q_rest"])
    else:
        o.set_material(mat["k_stretch"].astype(np.float32), mat["q_rest"].astype(np.float32))
    return o, mat


def subset(scene, strands):
    """The strands [0, strands) of a scene as a scene of their own."""
    nv = int(scene.strand_offsets[strands])
    return scenes.Scene(f"{scene.name}-first{strands}", scene.positions[:nv].copy(),
                        scene.strand_offsets[:strands + 1].copy(), material=scene.material)


def fp64_loss(sc, params, young, kappa, dt, nsteps, targets, threads=1):
    """Per-strand loss sum |x_tip - target|^2 of the FP64 oracle rollout."""
    o, _ = _sim(sc, params, young, kappa, 64, threads)
    for _ in range(nsteps):
        o.advance(dt)
    tips = sc.strand_offsets[1:] - 1
    x = o.get_pos_vel(f64=True)[0][tips]
    return np.sum((x - targets) ** 2, axis=1)


def strand_gradient(sc, params, young, kappa, dt, nsteps, K, targets, prec, threads=1):
    """Loss and per-strand (dL/dE, dL/dkappa) of the oracle in precision prec."""
    o, mat = _sim(sc, params, young, kappa, prec, threads)
    tips = (sc.strand_offsets[1:] - 1).astype(np.int32)
    loss, ksb, qrb = o.gradient(dt, nsteps, K, tips, targets)
    gE, gk = materials.strand_gradient(mat, sc.strand_offsets, ksb, qrb)
    return loss, gE, gk


def clamp_branches(sc, params, young, kappa, dt, nsteps, prec, threads=1):
    """Branch of the adaptive initial guess's clamp per substep and vertex."""
    o, _ = _sim(sc, params, young, kappa, prec, threads)
    o.log_clamp(True)
    for _ in range(nsteps):
        o.advance(dt)
    return o.log_clamp(False)


def central_fd(sc, params, young, kappa, dt, nsteps, targets, strand, which, h=1e-6):
    """Central finite difference of strand `strand`'s loss in its E
    (which = 0) or kappa (which = 1), step h relative, FP64 oracle; the
    strand is rerun alone."""
    one = subset_range(sc, strand, strand + 1)
    base = [np.array([young[strand]], np.float64), np.array([kappa[strand]], np.float64)]
    step = h * base[which][0]
    up = [b.copy() for b in base]
    dn = [b.copy() for b in base]
    up[which][0] += step
    dn[which][0] -= step
    tg = targets[strand:strand + 1]
    lu = fp64_loss(one, params, up[0], up[1], dt, nsteps, tg)[0]
    ld = fp64_loss(one, params, dn[0], dn[1], dt, nsteps, tg)[0]
    return (lu - ld) / (2 * step)


def subset_range(scene, s0, s1):
    """The strands [s0, s1) of a scene as a scene of their own."""
    a, b = int(scene.strand_offsets[s0]), int(scene.strand_offsets[s1])
    return scenes.Scene(f"{scene.name}-{s0}-{s1}", scene.positions[a:b].copy(),
                        (scene.strand_offsets[s0:s1 + 1] - a).astype(np.int32), material=scene.material)


def check(scene, params, young, kappa, dt, nsteps, K, targets, gpu_gE, gpu_gk, strands=16, fd_strands=4,
          threads=1, fd_steps=(1e-6, 1e-4)):
    """Compares a device gradient (per strand, whole groom) with the oracle on
    the first `strands` strands.  The FP64 gradient is pinned by central
    differences on the first `fd_strands`, at each relative step of
    `fd_steps` (at 1e-6 the FD of the smallest gradients carries FP64
    rounding noise of ~1e-3).  Returns a dict of the figures."""
    sc = subset(scene, strands)
    tg = np.asarray(targets, np.float64)[:strands]
    E = np.asarray(young, np.float64)[:strands]
    k = np.asarray(kappa, np.float64)[:strands]
    _, gE32, gk32 = strand_gradient(sc, params, E, k, dt, nsteps, K, tg, 32, threads)
    _, gE64, gk64 = strand_gradient(sc, params, E, k, dt, nsteps, K, tg, 64, threads)
    fd = {}
    for h in fd_steps:
        fd[h] = (np.array([central_fd(scene, params, young, kappa, dt, nsteps, targets, s, 0, h)
                           for s in range(fd_strands)]),
                 np.array([central_fd(scene, params, young, kappa, dt, nsteps, targets, s, 1, h)
                           for s in range(fd_strands)]))
    b32 = clamp_branches(sc, params, E, k, dt, nsteps, 32, threads)
    b64 = clamp_branches(sc, params, E, k, dt, nsteps, 64, threads)
    # positions: FP32 rollout noise against the FP64 displacement
    o32, _ = _sim(sc, params, E, k, 32, threads)
    o64, _ = _sim(sc, params, E, k, 64, threads)
    for _ in range(nsteps):
        o32.advance(dt)
        o64.advance(dt)
    x32 = o32.get_pos_vel()[0].astype(np.float64)
    x64 = o64.get_pos_vel(f64=True)[0]
    disp = np.abs(x64 - sc.positions).max()
    gE = np.asarray(gpu_gE, np.float64)[:strands]
    gk = np.asarray(gpu_gk, np.float64)[:strands]

    def rel(a, b):
        return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)

    out = {
        "strands_checked": int(strands),
        # the per-vertex adjoints are bit-identical; the per-strand sums of
        # their chain rule differ only in float64 summation order
        "device_vs_fp32_oracle_max_rel": float(max(rel(gE, gE32).max(), rel(gk, gk32).max())),
        "fp64_adjoint_vs_fp64_fd_max_rel": {
            f"h={h:g}": {"dE": float(rel(gE64[:fd_strands], fE).max()),
                         "dkappa": float(rel(gk64[:fd_strands], fk).max())} for h, (fE, fk) in fd.items()},
        # per strand, the FD step whose difference agrees best: at h = 1e-6
        # the smallest gradients carry FP64 rounding noise, at h = 1e-4 a
        # clamp branch of the adaptive initial guess can switch inside +-h
        "fp64_adjoint_vs_fp64_fd_best_step_max_rel": {
            "dE": float(np.max(np.min([rel(gE64[:fd_strands], fE) for fE, _ in fd.values()], axis=0))),
            "dkappa": float(np.max(np.min([rel(gk64[:fd_strands], fk) for _, fk in fd.values()], axis=0)))},
        "device_vs_fp64_gradient_rel": {
            "dE_median": float(np.median(rel(gE, gE64))), "dE_max": float(rel(gE, gE64).max()),
            "dkappa_median": float(np.median(rel(gk, gk64))), "dkappa_max": float(rel(gk, gk64).max())},
        "device_vs_fp64_fd_max_rel": {
            f"h={h:g}": {"dE": float(rel(gE[:fd_strands], fE).max()),
                         "dkappa": float(rel(gk[:fd_strands], fk).max())} for h, (fE, fk) in fd.items()},
        "clamp_branch_flips_fp32_vs_fp64": int(np.count_nonzero(b32 != b64)),
        "clamp_branches_logged": int(b32.size),
        "fp32_position_noise_over_fp64_displacement": float(np.abs(x32 - x64).max() / max(disp, 1e-300)),
        "fp64_displacement_m": float(disp),
    }
    return out
