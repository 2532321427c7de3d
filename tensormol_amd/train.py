"""Force-matching training step for the BP energy networks.

The reference trains energies and energy gradients together with
loss_op_EandG (TFNetworks/TFMolInstanceDirect.py:4889-4903) in
train_step_EandG (:5288-5348); the force term needs the derivative of
tf.gradients(E, xyz) (:5761) with respect to the weights, i.e. a second-order
graph.  Here that graph is split along its natural seam:

  G = descriptors(x)                     HIP forward kernel (not differentiated in W)
  eps_i = MLP_e(G_i; W)                  torch, FP32 or BF16 operands / FP32 accumulation
  J_i = d eps_i / d G_i                  torch autograd with create_graph
  dE/dx = (dG/dx)^T J                    HIP adjoint kernels (DescriptorAdjoint.forward)
  L(E, F = -dE/dx)
  dL/dJ = (dG/dx) . dL/d(dE/dx)          HIP tangent kernel (DescriptorAdjoint.backward)
  dL/dW                                  torch double backward through J

so the descriptor Jacobian is never stored: the adjoint and its transpose are
recomputed from the neighbour lists.  The loss is
L = sum_b w_e[b] (E_b - E*_b)^2 + sum_i w_f[b(i)] |F_i - F*_i|^2 (F = -dE/dx,
Ha/A); loss_op_EandG is w_e = EnergyScalar/2 (maxatom/natom_b)^2, w_f =
GradScalar/2 (maxatom/natom_b)^2, and MSE(E) + 0.5 MSE(F) is w_e = 1/B,
w_f = 0.5/(3N) (the default).
"""
from dataclasses import dataclass

import numpy as np
import torch
from torch.autograd.function import once_differentiable

from .engine import BPEEEngine


class DescriptorAdjoint(torch.autograd.Function):
    """dE/dx = (dG/dx)^T J on the HIP adjoint kernels; its derivative with
    respect to J, contracted with the incoming gradient, is the HIP tangent."""

    @staticmethod
    def forward(ctx, J, eng, ws, pos):
        grad = torch.empty(ws.N, 3, dtype=torch.float64, device=pos.device)
        eng.descriptor_vjp(ws, pos, J.detach().contiguous(), grad)
        ctx.eng, ctx.ws, ctx.pos = eng, ws, pos
        return grad

    @staticmethod
    @once_differentiable
    def backward(ctx, gg):
        V = torch.empty(ctx.ws.N, ctx.eng.D, dtype=torch.float32, device=gg.device)
        ctx.eng.descriptor_jvp(ctx.ws, ctx.pos, gg.double().contiguous(), V)
        return V, None, None, None


@dataclass
class TrainBatch:
    ws: object
    N: int
    B: int
    idx: dict              # element -> device index tensor of its atoms
    mol: torch.Tensor      # device int64 [N]
    E_ref: torch.Tensor    # float64 [B] (Ha)
    F_ref: torch.Tensor    # float64 [N, 3] (Ha/A)
    w_e: torch.Tensor      # float64 [B]
    w_f: torch.Tensor      # float64 [B]


class ForceMatchingTrainer:
    """Energy + force matching of a BPEEModel's energy networks.

    mlp_dtype "fp32" or "bf16" (GEMM operands rounded to BF16, products and
    accumulation in FP32; activations, their derivatives and all gradients in
    FP32).  The rounded operands go through FP32 GEMMs: on this software
    stack torch's BF16 GEMMs stall the host for milliseconds per call, and the
    step is bound by launch latency, not by GEMM throughput."""

    def __init__(self, model, device="cuda", mlp_dtype="fp32", lr=1e-4, ldr=128, lda=48,
                 capturable=False):
        """capturable: keep the optimizer's state on the device so that a whole
        step can be captured into a HIP graph (capture())."""
        if mlp_dtype not in ("fp32", "bf16"):
            raise ValueError(f"mlp_dtype {mlp_dtype!r}: use 'fp32' or 'bf16'")
        self.model = model
        self.eng = BPEEEngine(model, device=device, precision="fp32", ldr=ldr, lda=lda)
        self.device = self.eng.device
        self.alpha = float(model.alpha)
        self.mlp_dtype = mlp_dtype
        f32 = dict(dtype=torch.float32, device=self.device)
        self.params = {e: [(torch.tensor(np.asarray(W), **f32).requires_grad_(),
                            torch.tensor(np.asarray(b), **f32).requires_grad_())
                           for (W, b) in model.energy_nets[e]] for e in model.eles}
        self.opt = torch.optim.Adam(self.parameters(), lr=lr, capturable=bool(capturable))

    def parameters(self):
        """Flat list: elements in model order, layers in order, (W, b) each."""
        return [t for e in self.model.eles for (W, b) in self.params[e] for t in (W, b)]

    def batch(self, Z, mol_start, E_ref, F_ref, w_e=None, w_f=None):
        Z = np.asarray(Z)
        ms = np.asarray(mol_start, dtype=np.int64)
        N, B = len(Z), len(ms) - 1
        ws = self.eng.prepare(Z, ms, None)
        dev = self.device
        idx = {e: torch.as_tensor(np.nonzero(Z == e)[0], device=dev) for e in self.model.eles
               if np.any(Z == e)}
        mol = torch.as_tensor(np.repeat(np.arange(B), np.diff(ms)), device=dev)
        f64 = dict(dtype=torch.float64, device=dev)
        w_e = np.full(B, 1.0 / B) if w_e is None else np.asarray(w_e, dtype=np.float64)
        w_f = np.full(B, 0.5 / (3 * N)) if w_f is None else np.asarray(w_f, dtype=np.float64)
        return TrainBatch(ws, N, B, idx, mol, torch.as_tensor(np.asarray(E_ref), **f64),
                          torch.as_tensor(np.asarray(F_ref), **f64), torch.as_tensor(w_e, **f64),
                          torch.as_tensor(w_f, **f64))

    @staticmethod
    def _bf16(t):
        """Round to BF16 in the forward pass, identity in the backward pass."""
        return t + (t.to(torch.bfloat16).float() - t).detach()

    def _mlp(self, h, layers):
        a = self.alpha
        for k, (W, b) in enumerate(layers):
            if self.mlp_dtype == "bf16":
                # BF16 operands: their products are exact in FP32, so an FP32 GEMM
                # of the rounded operands is the arithmetic of a BF16 MFMA GEMM
                # with FP32 accumulation (up to summation order)
                z = self._bf16(h) @ self._bf16(W) + b
            else:
                z = h @ W + b
            if k < len(layers) - 1:
                h = torch.logaddexp(torch.zeros_like(z), a * z) / a   # log(1 + e^{a z}) / a
            else:
                h = z
        return h[:, 0]

    def loss(self, tb, pos, check=True):
        """(L, E, F) at positions pos (cuda float64 [N, 3]); differentiable in
        the network weights.  check reads the neighbour-list flags on the host
        (skipped inside a graph capture)."""
        eng, ws = self.eng, tb.ws
        eng.build_lists(ws, pos)
        if check:
            eng.check_flags(ws)
        eng.descriptors(ws, pos)
        G = ws.G[:tb.N].detach().clone().requires_grad_(True)
        eps = torch.zeros(tb.N, dtype=torch.float32, device=self.device)
        for e, idx in tb.idx.items():
            eps = eps.index_put((idx,), self._mlp(G[idx], self.params[e]))
        E = torch.zeros(tb.B, dtype=torch.float64, device=self.device).index_add(
            0, tb.mol, eps.double())
        (J,) = torch.autograd.grad(eps.sum(), G, create_graph=True)
        F = -DescriptorAdjoint.apply(J, eng, ws, pos)
        L = (tb.w_e * (E - tb.E_ref) ** 2).sum() + (tb.w_f[tb.mol][:, None] * (F - tb.F_ref) ** 2).sum()
        return L, E, F

    def gradients(self, tb, pos):
        L, E, F = self.loss(tb, pos)
        return L, torch.autograd.grad(L, self.parameters())

    def step(self, tb, pos, check=True):
        """One Adam step; returns the loss before the update."""
        self.opt.zero_grad(set_to_none=True)
        L, _, _ = self.loss(tb, pos, check=check)
        L.backward()
        self.opt.step()
        return L

    def capture(self, tb, pos, warmup=3):
        """Capture one whole step (lists, descriptors, MLP, double backward,
        Adam update) for a fixed batch and positions into a HIP graph; returns
        (graph, loss) where each graph.replay() performs one more step and
        leaves its loss (before the update) in `loss`.  Needs capturable=True.
        The `warmup` eager steps before the capture are training steps too."""
        if not self.opt.param_groups[0].get("capturable", False):
            raise RuntimeError("capture() needs ForceMatchingTrainer(..., capturable=True)")
        cur = torch.cuda.current_stream(self.device)
        side = torch.cuda.Stream(self.device)
        side.wait_stream(cur)
        with torch.cuda.stream(side):
            for _ in range(warmup):
                self.step(tb, pos)
        cur.wait_stream(side)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        self.opt.zero_grad(set_to_none=True)
        with torch.cuda.graph(g):
            L, _, _ = self.loss(tb, pos, check=False)
            L.backward()
            self.opt.step()
        return g, L
