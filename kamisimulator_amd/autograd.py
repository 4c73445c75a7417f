"""PyTorch binding of the differentiable step: a rollout of a strand scene as
a ``torch.autograd.Function`` whose backward is the HIP adjoint
(kscr_adjoint_forward / kscr_adjoint_backward, include/kscr.h).

Parameters are per strand: Young's modulus E [GPa] and a rest-curvature scale
kappa (the rest rotor of every segment raised to the power kappa, so kappa = 1
is the groom's own shape).  They map to the per-vertex kStretch and qRest the
reference's Sim::setYoung computes (YBSimulator.cu:300-320) with torch ops in
float64, so autograd carries the chain rule from the per-vertex adjoints the
HIP library returns back to (E, kappa).

The reference has no differentiable path; this binding is test and research
plumbing around the product kernels, not part of the solver interface.
"""
import numpy as np
import torch

from . import scenes
from ._lib import Sim, default_params


def _rotor_power(kappa, r):
    """kappa * r for unit rotors r[..., 4] (x, y, z, w): same angle axis,
    angle scaled by kappa (KE/includes/modules/Rotor.h:221-228)."""
    qv = r[..., :3]
    w = r[..., 3].clamp(-1.0, 1.0)
    qn = torch.linalg.norm(qv, dim=-1)
    ident = (w == 1.0) | (qn == 0)
    safe = torch.where(ident, torch.ones_like(qn), qn)
    axis = qv / safe[..., None]
    alpha = torch.acos(torch.where(ident, torch.zeros_like(w), w))
    na = kappa * alpha
    val = torch.cat([axis * torch.sin(na)[..., None], torch.cos(na)[..., None]], dim=-1)
    unit = torch.tensor([0.0, 0.0, 0.0, 1.0], dtype=r.dtype, device=r.device)
    return torch.where(ident[..., None], unit, val)


def strand_material(rest_dir, l_rest, counts, young_gpa, kappa, radius, deviation=1.0):
    """Per-strand E [GPa] and kappa -> per-vertex (kStretch, qRest), the map of
    Sim::setYoung (YBSimulator.cu:300-320) with the rest rotors scaled by kappa:
        kStretch_i = E pi mmrad^2,  bt_i = kStretch_i deviation mmrad^2 / (4 l_i^2),
        qRest_i = bt_i * (kappa * r_i),  mmrad = radius * 1000.
    rest_dir: unit rest rotors (n, 4); l_rest (n); counts: vertices per strand."""
    mmrad = float(radius) * 1000.0
    E = torch.repeat_interleave(young_gpa, counts)
    K = torch.repeat_interleave(kappa, counts)
    ks = E * (np.pi * mmrad * mmrad)
    bt = ks * deviation * mmrad * mmrad * 0.25 / (l_rest * l_rest)
    return ks, bt[:, None] * _rotor_power(K, rest_dir)


class _SimHolder:
    """Owns the handle of one rollout.  Its device buffers (state, checkpoints
    and tape, which can take GBs) are released when the autograd graph that
    holds the rollout is freed: after the backward pass unless the graph is
    retained, or when the outputs are dropped without one."""

    def __init__(self, sim):
        self.sim = sim

    def __del__(self):
        sim = self.sim
        self.sim = None
        if sim is not None:
            try:
                sim.close()
            except Exception:
                pass


class _RolloutFn(torch.autograd.Function):
    """(kStretch[n], qRest[n, 4]) -> (x[n, 3], v[n, 3]) after the rollout."""

    @staticmethod
    def forward(ctx, k_stretch, q_rest, rollout):
        sim = rollout.fresh_sim()
        sim.set_material(k_stretch.detach().cpu().numpy(), q_rest.detach().cpu().numpy())
        sim.adjoint_forward(rollout.dt, rollout.nsteps, rollout.checkpoint_every)
        x, v = sim.get_pos_vel()
        ctx.holder = _SimHolder(sim)
        ctx.dev = k_stretch.device
        return torch.from_numpy(x).to(ctx.dev), torch.from_numpy(v).to(ctx.dev)

    @staticmethod
    def backward(ctx, gx, gv):
        # kscr_adjoint_backward re-records every segment it needs from the
        # checkpoints and leaves the state at the end of the rollout, so a
        # retained graph can run it again
        xb = None if gx is None else gx.detach().float().cpu().numpy()
        vb = None if gv is None else gv.detach().float().cpu().numpy()
        ksb, qrb = ctx.holder.sim.adjoint_backward(xb, vb)
        return torch.from_numpy(ksb).to(ctx.dev), torch.from_numpy(qrb).to(ctx.dev), None


class Rollout:
    """nsteps steps of dt of one scene from its rest state, differentiable in
    the per-strand (E, kappa).  Every call starts from a fresh handle, so the
    rollout is a pure function of its parameters."""

    def __init__(self, scene, dt, nsteps, checkpoint_every, device=-1):
        self.scene = scene
        self.dt = np.float32(dt)
        self.nsteps = int(nsteps)
        self.checkpoint_every = int(checkpoint_every)
        self.device = device
        self.params = scenes.make_params(default_params, scene.material)
        probe = self.fresh_sim()
        rest = probe.get_rest()
        probe.close()
        qr = torch.from_numpy(rest["q_rest"].astype(np.float64))
        nrm = torch.linalg.norm(qr, dim=1, keepdim=True)
        unit = torch.tensor([0.0, 0.0, 0.0, 1.0], dtype=torch.float64)
        self.rest_dir = torch.where(nrm > 0, qr / torch.where(nrm > 0, nrm, torch.ones_like(nrm)), unit)
        self.l_rest = torch.from_numpy(rest["l_rest"].astype(np.float64))
        self.counts = torch.from_numpy(np.diff(scene.strand_offsets).astype(np.int64))
        self.n = int(scene.num_verts)

    def fresh_sim(self):
        sim = Sim(self.params, self.scene.positions, self.scene.strand_offsets, device=self.device)
        scenes.setup_sim(sim, self.scene)
        return sim

    def material(self, young_gpa, kappa):
        """Per-strand E [GPa] and kappa (float64 tensors) -> per-vertex
        (kStretch, qRest) in float64, differentiable."""
        return strand_material(self.rest_dir, self.l_rest, self.counts, young_gpa, kappa, self.params.radius,
                               self.scene.material.deviation)

    def __call__(self, young_gpa, kappa):
        """Final positions and velocities (float32 tensors [n, 3])."""
        ks, qr = self.material(young_gpa, kappa)
        return _RolloutFn.apply(ks.float(), qr.float(), self)
