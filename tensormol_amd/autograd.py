"""BP+EE energy as a torch.autograd.Function.

The reference obtains forces with tf.gradients(Etotal, xyzs)
(TFNetworks/TFMolInstanceDirect.py:5761) and returns -JOULEPERHARTREE times
that gradient (TFNetworks/TFMolManage.py:1320).  Here the forward pass runs
the engine once (descriptors, networks, Coulomb and the hand-written adjoint
kernels together) and keeps dE/dx; backward scales it by the incoming
per-molecule gradient.  Second derivatives are not provided.
"""
import torch
from torch.autograd.function import once_differentiable


class BPEEEnergy(torch.autograd.Function):
    """E[b] (Ha) of every molecule of the workspace's batch at positions pos
    (cuda float64 [N, 3], A); d E[b] / d pos from the adjoint kernels."""

    @staticmethod
    def forward(ctx, pos, engine, ws):
        p = pos.detach().contiguous()
        engine.run(ws, p, forces=True)
        E = engine.energy_tensor(ws).clone()
        ctx.save_for_backward(ws.grad[:ws.N].clone(), ws.mol[:ws.N].long())
        return E

    @staticmethod
    @once_differentiable
    def backward(ctx, gE):
        grad, mol = ctx.saved_tensors
        return gE[mol].unsqueeze(1) * grad, None, None


def energy(engine, ws, pos):
    """Per-molecule total energy with autograd support for pos."""
    if pos.dtype != torch.float64 or not pos.is_cuda:
        raise TypeError("pos must be a cuda float64 tensor")
    if pos.shape != (ws.N, 3):
        raise ValueError(f"pos has shape {tuple(pos.shape)}, the batch has {ws.N} atoms")
    return BPEEEnergy.apply(pos, engine, ws)
