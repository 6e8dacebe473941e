"""Differentiable AVBD rollouts as a torch.autograd.Function.

`rollout(engine, scene, v0, mass, friction, nsteps)` steps `scene` for
`nsteps` fixed steps on the GPU (oavbd_diff_forward, a checkpoint every
`every` steps) with the bodies' initial velocities, masses and friction
coefficients taken from the given tensors, and returns the final body state
(n x 13: position, rotation, velocity, angular velocity).  Its backward is
the library's hand-written adjoint of the step (oavbd_diff_backward): each
checkpoint interval is re-run on the device while its steps are recorded,
then differentiated backwards through each step's fixed active set.
Gradients flow to v0 (n x 3), mass (n; zero for static bodies) and friction
(n; a pair's friction is the larger of its bodies' coefficients, and that
body receives the pair's gradient).  This is a thin binding: the tensors
cross to the host once per call; all stepping and differentiation run in
liboavbd.so.
"""
from __future__ import annotations

import numpy as np
import torch

from .avbd import AvbdEngine


class AvbdRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, mass, friction, engine, scene, nsteps, every):
        s = dict(scene)
        s["vel"] = np.ascontiguousarray(v0.detach().cpu().numpy(), dtype=np.float32)
        s["mass"] = np.ascontiguousarray(mass.detach().cpu().numpy(), dtype=np.float32)
        s["friction"] = np.ascontiguousarray(friction.detach().cpu().numpy(), dtype=np.float32)
        engine.load(s, scene_ids=s.get("scene_id"))
        engine.diff_forward(int(nsteps), int(every))
        ctx.engine = engine
        ctx.device = v0.device
        return torch.from_numpy(engine.state()).to(v0.device)

    @staticmethod
    def backward(ctx, g_final):
        d0, dm, df = ctx.engine.diff_backward(g_final.detach().cpu().numpy())
        dev = ctx.device
        return (torch.from_numpy(np.ascontiguousarray(d0[:, 7:10])).to(dev), torch.from_numpy(dm).to(dev),
                torch.from_numpy(df).to(dev), None, None, None, None)


def rollout(engine: AvbdEngine, scene, v0, mass, friction, nsteps, every=10):
    """Final body state (n x 13) after `nsteps` steps; differentiable with
    respect to v0 (n x 3), mass (n) and friction (n)."""
    return AvbdRollout.apply(v0, mass, friction, engine, scene, nsteps, every)


def centroid_loss(final, scene_id, dynamic, targets):
    """Sum over scenes of |centroid of the scene's dynamic bodies - target|^2."""
    sid = torch.as_tensor(scene_id.astype(np.int64), device=final.device)
    dyn = torch.as_tensor(dynamic, device=final.device)
    ns = targets.shape[0]
    pos = final[:, 0:3] * dyn[:, None].to(final.dtype)
    sums = torch.zeros(ns, 3, dtype=final.dtype, device=final.device).index_add_(0, sid, pos)
    cnt = torch.zeros(ns, dtype=final.dtype, device=final.device).index_add_(0, sid, dyn.to(final.dtype))
    c = sums / cnt[:, None]
    return ((c - targets) ** 2).sum()


__all__ = ["AvbdRollout", "rollout", "centroid_loss"]
