"""PyTorch autograd over the engine's recorded trajectories.

`simulate(x0, v0, m, nsteps, engine)` runs `nsteps` Barnes-Hut frames on the
GPU engine (bhx_record) and returns (x_T, v_T) as tensors; their backward is
the engine's reverse pass (bhx_adjoint): gradients reach x0, v0 and m with the
walk's decisions frozen at the recorded frames.  PyTorch only carries the
tensors; every frame and every adjoint runs in libbhx.

An engine holds one recording at a time: call backward before the next
simulate on the same engine.  A backward pass whose recording has since been
replaced (another simulate, record or set_bodies on that engine) raises
instead of returning the gradients of the other trajectory.
"""
import os
import sys

import torch

if "bhx" not in sys.modules:
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from bhx import Engine  # noqa: E402


class _Trajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, m, engine, nsteps, checkpoint_every, init_acc):
        dev = x0.device
        engine.set_bodies(x0.detach().to("cpu", torch.float64).numpy(),
                          v0.detach().to("cpu", torch.float64).numpy(),
                          m.detach().to("cpu", torch.float64).numpy())
        engine.record(nsteps, checkpoint_every, init_acc)
        P, V, _, _ = engine.recorded_bodies()
        ctx.engine = engine
        ctx.generation = engine.record_generation
        ctx.devices = (x0.device, v0.device, m.device)
        return torch.from_numpy(P).to(dev), torch.from_numpy(V).to(dev)

    @staticmethod
    def backward(ctx, gx, gv):
        if ctx.engine.record_generation != ctx.generation:
            raise RuntimeError("the engine's recording was replaced after this simulate(): "
                               "call backward before simulating again on the same engine")

        def host(g):
            return None if g is None else g.detach().to("cpu", torch.float64).contiguous().numpy()
        gx0, gv0, _, gm = ctx.engine.adjoint(host(gx), host(gv))
        dx, dv, dm = ctx.devices
        return (torch.from_numpy(gx0).to(dx), torch.from_numpy(gv0).to(dv), torch.from_numpy(gm).to(dm),
                None, None, None, None)


def make_engine(n, **cfg):
    """An engine configured for differentiation: group walk, intended moments, monopole."""
    base = dict(theta=0.5, eps=0.01, dt=1e-3, leaf_capacity=32, mode=1, precision=1, quadrupole=0,
                permute_acc=1, group_walk=1)
    base.update(cfg)
    return Engine(n, **base)


def simulate(x0, v0, m, nsteps, engine, checkpoint_every=100, init_acc=True):
    """(x_T, v_T) after nsteps frames from (x0, v0, m), differentiable in all three.
    Shapes (N, 3), (N, 3), (N,); any device (the engine keeps its own copies in HBM)."""
    if x0.shape != v0.shape or x0.dim() != 2 or x0.shape[1] != 3 or m.shape != (x0.shape[0],):
        raise ValueError("expected x0, v0 of shape (N, 3) and m of shape (N,)")
    return _Trajectory.apply(x0, v0, m, engine, int(nsteps), int(checkpoint_every), bool(init_acc))


__all__ = ["simulate", "make_engine"]
