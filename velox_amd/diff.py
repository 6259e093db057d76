"""Differentiable rollouts of batched worlds on the device, behind a PyTorch
autograd function.

    rollout = Rollout(Scene("diff", 1024), steps=120, every=8)
    loss = RolloutLoss.apply(v0, w0, friction, mass, rollout)   # (W,)
    loss.sum().backward()

v0, w0 are (N, 3) tensors of initial linear and angular velocities,
friction and mass (N,) tensors, N the bodies of the scene (static bodies'
entries are ignored).  The forward is the device step over the rollout with
a checkpoint kept on the device every `every` steps; the backward is the
library's hand-written adjoint kernels (vxhip_diff_backward) over the
frozen active set of each step.  The tensors cross the C ABI as host arrays;
gradients come back on the loss's device.  One forward/backward pair at a
time per Rollout: a new forward replaces the checkpoints of the last one.
"""
import numpy as np
import torch

from .device import Context

SEED_TARGETS = 5005


def default_targets(scene, seed=SEED_TARGETS):
    """A seeded target per world: its initial centre of mass moved by
    (U(1, 3), 0, U(-1, 1)) metres."""
    rng = np.random.default_rng(seed)
    b = scene.bodies
    dyn = b["motionType"] == 2
    out = np.zeros((scene.worlds, 3), np.float32)
    mass = np.where(dyn, 1.0 / np.where(b["invMass"] > 0, b["invMass"], 1.0), 0.0)
    for w in range(scene.worlds):
        sel = (scene.world_of == w) & dyn
        com = (b["position"][sel] * mass[sel, None]).sum(0) / mass[sel].sum()
        out[w] = com + np.array([rng.uniform(1.0, 3.0), 0.0, rng.uniform(-1.0, 1.0)], np.float32)
    return out


class Rollout:
    """A batched scene, rollout length, checkpoint interval and targets, on
    one device context."""

    def __init__(self, scene, steps=120, every=8, targets=None, context=None):
        self.scene = scene
        self.steps, self.every = int(steps), int(every)
        self.targets = np.ascontiguousarray(default_targets(scene) if targets is None else targets,
                                            np.float32).reshape(-1, 3)
        self.ctx = context or Context()
        b = scene.bodies
        self.dynamic = b["motionType"] == 2
        im = np.where(b["invMass"] > 0, b["invMass"], 1.0)
        self.kappa = (b["invInertia"][:, 0] / im).astype(np.float32)  # inverse inertia times mass
        self.last_stats = None

    def initial(self):
        """The scene's own parameters: v0 (N, 3), w0 (N, 3), friction (N,), mass (N,)."""
        b = self.scene.bodies
        mass = np.where(self.dynamic, 1.0 / np.where(b["invMass"] > 0, b["invMass"], 1.0), 0.0).astype(np.float32)
        return (b["velocity"].astype(np.float32), b["angularVelocity"].astype(np.float32),
                b["friction"].astype(np.float32), mass)

    def bodies_with(self, v0, w0, friction, mass):
        """The scene's body records with these parameters (dynamic bodies)."""
        b = self.scene.bodies.copy()
        d = self.dynamic
        b["velocity"][d] = v0[d]
        b["angularVelocity"][d] = w0[d]
        b["friction"][d] = friction[d]
        m = np.asarray(mass, np.float32)[d]
        b["invMass"][d] = np.float32(1.0) / m
        b["invInertia"][d] = (self.kappa[d] / m)[:, None]
        return b

    def load(self, v0, w0, friction, mass):
        sc = self.scene
        self.ctx.load(self.bodies_with(v0, w0, friction, mass), sc.soup, gravity=sc.gravity, substeps=sc.substeps)
        if sc.worlds > 1:
            self.ctx.set_worlds(sc.world_of, sc.worlds)
        return self.ctx

    def forward(self, v0, w0, friction, mass):
        self.load(v0, w0, friction, mass)
        return self.ctx.diff_forward(self.scene.dt, self.steps, self.every, self.targets)

    def backward(self, loss_weight):
        gv, gw, gf, gm, st = self.ctx.diff_backward(loss_weight)
        self.last_stats = st
        return gv, gw, gf, gm


def _host(t):
    return t.detach().to("cpu", torch.float32).contiguous().numpy()


class RolloutLoss(torch.autograd.Function):
    """Per-world loss |centre of mass after the rollout - target|^2."""

    @staticmethod
    def forward(ctx, v0, w0, friction, mass, rollout):
        loss = rollout.forward(_host(v0), _host(w0), _host(friction), _host(mass))
        ctx.rollout = rollout
        return torch.from_numpy(loss).to(v0.device)

    @staticmethod
    def backward(ctx, grad_loss):
        gv, gw, gf, gm = ctx.rollout.backward(_host(grad_loss))
        dev = grad_loss.device
        return (torch.from_numpy(gv).to(dev), torch.from_numpy(gw).to(dev), torch.from_numpy(gf).to(dev),
                torch.from_numpy(gm).to(dev), None)
