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
recomputed from the neighbour lists.

With mlp_dtype "bf16" the MLP half runs on hand-written MFMA kernels
(tm_gemm_bf16, BF16 operands, FP32 accumulation) instead of torch autograd:
per element, with rows r of its atoms, layers z_l = h_{l-1} W_l + b_l,
h_l = act(z_l), s_l = act'(z_l) = sigmoid(a z_l), y = h_3 w_4 + b_4,
  forward:    3 GEMMs with the bias / softplus epilogue, 1 for y;
  J = dy/dG:  zd_3 = s_3 * w_4, hd_{l-1} = zd_l W_l^T, zd_{l-1} = s_{l-1} * hd_{l-1},
              J = zd_1 W_1^T (3 GEMMs);
  then F = -(dG/dx)^T J, the loss, ybar = dLoss/dy and V = dLoss/dJ (HIP
  adjoint and tangent of the descriptors);
  tangent:    zt_l = ht_{l-1} W_l, ht_l = s_l * zt_l, ht_0 = V (3 GEMMs);
  reverse:    zb_3 = s_3 * ybar w_4 + a s_3 (1 - s_3) zt_3 w_4, hb_{l-1} = zb_l W_l^T,
              zb_{l-1} = s_{l-1} * hb_{l-1} + a s (1 - s) zt_{l-1} hd_{l-1} (2 GEMMs);
  gradients:  dW_l, db_l = [h_{l-1}, 1; ht_{l-1}, 0]^T [zb_l; zd_l], dw_4, db_4 =
              [h_3, 1; ht_3, 0]^T [ybar; 1] (4 GEMMs over the stacked rows).
That is the reverse pass of Phi = sum_r (ybar_r y_r + J_r . V_r), whose weight
gradient is the loss's (V held fixed).  The loss is
L = sum_b w_e[b] (E_b - E*_b)^2 + sum_i w_f[b(i)] |F_i - F*_i|^2 (F = -dE/dx,
Ha/A); loss_op_EandG is w_e = EnergyScalar/2 (maxatom/natom_b)^2, w_f =
GradScalar/2 (maxatom/natom_b)^2, and MSE(E) + 0.5 MSE(F) is w_e = 1/B,
w_f = 0.5/(3N) (the default).
"""
from dataclasses import dataclass

import numpy as np
import torch
from torch.autograd.function import once_differentiable

import ctypes as _ctypes

from . import _lib
from .engine import BPEEEngine

_lib_ptr = _lib.ptr


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
        step can be captured into a HIP graph (capture()).
        mlp_dtype: "fp32" (torch, FP32 GEMMs), "bf16" (hand-written MFMA kernels,
        BF16 operands) or "bf16_torch" (torch with BF16-rounded operands in FP32
        GEMMs, the same arithmetic up to the rounding of the backward GEMMs'
        gradient operands and the summation order)."""
        if mlp_dtype not in ("fp32", "bf16", "bf16_torch"):
            raise ValueError(f"mlp_dtype {mlp_dtype!r}: use 'fp32', 'bf16' or 'bf16_torch'")
        self.model = model
        self.eng = BPEEEngine(model, device=device, precision="fp32", ldr=ldr, lda=lda)
        self.device = self.eng.device
        self.alpha = float(model.alpha)
        self.mlp_dtype = mlp_dtype
        # a list to collect (start event, end event, flops) of every GEMM launch
        # of the MFMA path (eager steps only), or None
        self.gemm_timing = None
        f32 = dict(dtype=torch.float32, device=self.device)
        self.params = {e: [(torch.tensor(np.asarray(W), **f32).requires_grad_(),
                            torch.tensor(np.asarray(b), **f32).requires_grad_())
                           for (W, b) in model.energy_nets[e]] for e in model.eles}
        # one fused update kernel for every parameter on the GPU
        self.opt = torch.optim.Adam(self.parameters(), lr=lr, capturable=bool(capturable),
                                    fused=self.device.type == "cuda")

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
            if self.mlp_dtype == "bf16_torch":
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
        if self.mlp_dtype == "bf16":
            L, _, _, grads = self._mfma_step(tb, pos)
            return L, grads
        L, E, F = self.loss(tb, pos)
        return L, torch.autograd.grad(L, self.parameters())

    def step(self, tb, pos, check=True):
        """One Adam step; returns the loss before the update."""
        self.opt.zero_grad(set_to_none=True)
        if self.mlp_dtype == "bf16":
            L, _, _, grads = self._mfma_step(tb, pos, check=check, views=True)
            for p, g in zip(self.parameters(), grads):
                p.grad = g
        else:
            L, _, _ = self.loss(tb, pos, check=check)
            L.backward()
        self.opt.step()
        return L

    # ------------------------------------------------ MFMA path (mlp_dtype "bf16")
    def _mfma_buffers(self, tb):
        """Per-element device buffers of a batch (made once per batch).  Row
        stride ld = width + 16; the column after the last feature of the
        stacked inputs holds 1 for the primal rows and 0 for the tangent rows
        (the bias column of the weight-gradient GEMMs).  Arrays that only
        GEMMs read (stacked inputs X0, activations H, upstream rows ZZ, the
        weights' copies Wb, YBb) are BF16: the GEMMs round their operands to
        BF16 anyway, so storing them rounded gives the same products."""
        bufs = getattr(tb, "_mfma", None)
        if bufs is not None:
            return bufs
        dev, D = self.device, self.eng.D
        H = [int(W.shape[1]) for (W, _) in self.params[self.model.eles[0]]][:3]
        assert H[0] == H[1] == H[2], "the MFMA training path needs equal hidden widths"
        H = H[0]
        f32 = dict(dtype=torch.float32, device=dev)
        b16 = dict(dtype=torch.bfloat16, device=dev)
        ldx, ldh = D + 16, H + 16
        bufs = {}
        for e, idx in tb.idx.items():
            M = int(idx.numel())
            b = {"M": M, "idx": idx.to(torch.int32).contiguous(), "ldx": ldx, "ldh": ldh, "H": H}
            b["X0"] = torch.zeros(2 * M, ldx, **b16)
            b["X0"][:M, D] = 1.0
            for k in ("H1", "H2", "H3"):
                b[k] = torch.zeros(2 * M, ldh, **b16)
                b[k][:M, H] = 1.0
            for k in ("ZZ1", "ZZ2", "ZZ3"):
                b[k] = torch.zeros(2 * M, ldh, **b16)
            for k in ("S1", "S2", "S3", "ZT1", "ZT2", "ZT3", "HD1", "HD2"):
                b[k] = torch.zeros(M, ldh, **f32)
            b["YB"] = torch.zeros(M, 1, **f32)      # ybar (the last tangent layer's epilogue)
            b["YBb"] = torch.ones(2 * M, 1, **b16)  # [ybar; 1] (the w_4 gradient GEMM)
            b["Wb"] = [torch.zeros(W.shape, **b16) for (W, _) in self.params[e]]
            b["gW"] = [torch.zeros(D + 1, H, **f32), torch.zeros(H + 1, H, **f32),
                       torch.zeros(H + 1, H, **f32), torch.zeros(H + 1, 1, **f32)]
            bufs[e] = b
        # the weight-gradient GEMMs of every element run in one grouped launch
        # with a common K split (K = 2M rows); each keeps its own partials
        ns = max(1, min(8, (max(2 * b["M"] for b in bufs.values()) + 2047) // 2048))
        for b in bufs.values():
            b["nsplit"] = ns
            b["work"] = [torch.zeros(ns * (D + 1) * H, **f32), torch.zeros(ns * (H + 1) * H, **f32),
                         torch.zeros(ns * (H + 1) * H, **f32), torch.zeros(ns * (H + 1), **f32)]
        tb._mfma = bufs
        return bufs

    def _ggemm(self, groups, ta, tb_, epi=0, io=1, nsplit=1):
        """One tm_gemm_bf16_grouped call over dicts of tm_gemm_group fields
        (device addresses as ints; absent fields are NULL); io: the call's
        storage types (bit 0 BF16 operands, bits 1-3 BF16 C / aux / out2)."""
        arr = (_lib.GemmGroup * len(groups))(*[_lib.GemmGroup(**g) for g in groups])
        call = lambda: self.eng.lib.tm_gemm_bf16_grouped(
            len(groups), _ctypes.cast(arr, _ctypes.c_void_p), ta, tb_, epi, io, self.alpha,
            nsplit, self.eng._stream())
        if self.gemm_timing is None:
            call()
            return
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        call()
        ev[1].record()
        self.gemm_timing.append((ev[0], ev[1], sum(2 * g["M"] * g["N"] * g["K"] for g in groups)))

    def _grows(self, sets, scatter):
        """One tm_rows_grouped call over (M, cols, idx, in, ldi, out, ldo,
        out_bf16)."""
        arr = (_lib.RowsGroup * len(sets))(*[_lib.RowsGroup(*r) for r in sets])
        self.eng.lib.tm_rows_grouped(len(sets), _ctypes.cast(arr, _ctypes.c_void_p), scatter,
                                     self.eng._stream())

    def _cast(self, pairs):
        """One tm_cast_bf16_grouped call: each (FP32 tensor, BF16 tensor) pair."""
        arr = (_lib.CastGroup * len(pairs))(
            *[_lib.CastGroup(x.data_ptr(), y.data_ptr(), x.numel()) for x, y in pairs])
        self.eng.lib.tm_cast_bf16_grouped(len(pairs), _ctypes.cast(arr, _ctypes.c_void_p),
                                          self.eng._stream())

    def _mfma_step(self, tb, pos, check=True, views=False):
        """(L, E, F, gradients in parameters() order) of the BF16 MFMA path.
        Each GEMM stage runs for every element in one grouped launch, with
        the elementwise steps between them in the GEMM epilogues; the row
        gathers and scatters of every element also take one launch each.
        views=True returns the gradients as views of the batch's buffers (the
        next step overwrites them) instead of copies."""
        eng, ws = self.eng, tb.ws
        eng.build_lists(ws, pos)
        if check:
            eng.check_flags(ws)
        eng.descriptors(ws, pos)
        N, D = tb.N, eng.D
        bufs = self._mfma_buffers(tb)
        items = [(b, self.params[e]) for e, b in bufs.items()]
        H = items[0][0]["H"]
        ad = lambda t, rows=0, ld=0: t.data_ptr() + t.element_size() * rows * ld   # row `rows`
        f32 = dict(dtype=torch.float32, device=self.device)
        # every atom has an element of the model (prepare() checks), so the
        # per-element GEMMs below write every row of these through their row maps
        eps = torch.empty(N, 1, **f32)
        J = torch.empty(N, D, **f32)
        # this step's weights as BF16 (one launch), then the descriptors as BF16 rows
        self._cast([(prm[l][0], b["Wb"][l]) for b, prm in items for l in range(4)])
        self._grows([(b["M"], D, ad(b["idx"]), ad(ws.G), D, ad(b["X0"]), b["ldx"], 1)
                     for b, _ in items], 0)
        # forward: z_l = h_{l-1} W_l + b_l, h_l = act(z_l), s_l = act'(z_l); the
        # last layer also starts the J chain, zd_3 = s_3 w_4 (rows M.. of ZZ3)
        layers = (("X0", "H1", "S1", "ZT1"), ("H1", "H2", "S2", "ZT2"), ("H2", "H3", "S3", "ZT3"))
        for l, (hin, hout, sk, _) in enumerate(layers):
            grp = []
            for b, prm in items:
                M, ldh = b["M"], b["ldh"]
                ldp, Kp = (b["ldx"], D) if l == 0 else (ldh, H)
                g = dict(M=M, N=H, K=Kp, A=ad(b[hin]), lda=ldp, B=ad(b["Wb"][l]), ldb=H,
                         C=ad(b[hout]), ldc=ldh, bias=ad(prm[l][1]), aux=ad(b[sk]))
                if l == 2:
                    g.update(x1=ad(prm[3][0]), out2=ad(b["ZZ3"], M, ldh))
                grp.append(g)
            self._ggemm(grp, 0, 0, 1, io=1 | 2 | 8)   # H_l, zd_3 BF16; s_l FP32
        # atom energies y = h_3 w_4 + b_4, written to each atom's row of eps
        self._ggemm([dict(M=b["M"], N=1, K=H, A=ad(b["H3"]), lda=b["ldh"], B=ad(b["Wb"][3]),
                          ldb=1, C=ad(eps), ldc=1, bias=ad(prm[3][1]), crow=ad(b["idx"]))
                     for b, prm in items], 0, 0, 0)
        # J chain: hd_{l-1} = zd_l W_l^T, zd_{l-1} = s_{l-1} hd_{l-1}, J = zd_1 W_1^T
        for (l, zz_in, hd, sk, zz_out) in ((2, "ZZ3", "HD2", "S2", "ZZ2"),
                                           (1, "ZZ2", "HD1", "S1", "ZZ1")):
            self._ggemm([dict(M=b["M"], N=H, K=H, A=ad(b[zz_in], b["M"], b["ldh"]), lda=b["ldh"],
                              B=ad(b["Wb"][l]), ldb=H, C=ad(b[hd]), ldc=b["ldh"], s=ad(b[sk]),
                              aux=ad(b[zz_out], b["M"], b["ldh"])) for b, prm in items], 0, 1, 2,
                        io=1 | 4)   # hd FP32, zd BF16
        # J rows written to each atom's row of J
        self._ggemm([dict(M=b["M"], N=D, K=H, A=ad(b["ZZ1"], b["M"], b["ldh"]), lda=b["ldh"],
                          B=ad(b["Wb"][0]), ldb=H, C=ad(J), ldc=D, crow=ad(b["idx"]))
                     for b, prm in items], 0, 1, 0)
        f64 = dict(dtype=torch.float64, device=self.device)
        grad = torch.empty(N, 3, **f64)
        eng.descriptor_vjp(ws, pos, J, grad)
        # molecule energies, the loss, ybar = dLoss/dy and gg = dLoss/dgrad in one
        # deterministic launch pair
        E, part, Lt = torch.empty(tb.B, **f64), torch.empty(tb.B, **f64), torch.empty(1, **f64)
        ybar = torch.empty(N, 1, **f32)
        gg = torch.empty(N, 3, **f64)
        P = _lib_ptr
        eng.lib.tm_fm_loss(tb.B, P(ws.mol_start), P(eps), P(tb.E_ref), P(tb.w_e), P(grad),
                           P(tb.F_ref), P(tb.w_f), P(E), P(ybar), P(gg), P(part), P(Lt),
                           eng._stream())
        L = Lt[0]
        F = -grad
        V = torch.empty(N, D, **f32)
        eng.descriptor_jvp(ws, pos, gg, V)
        self._grows([(b["M"], D, ad(b["idx"]), ad(V), D, ad(b["X0"], b["M"], b["ldx"]), b["ldx"],
                      1) for b, _ in items] +
                    [(b["M"], 1, ad(b["idx"]), ad(ybar), 1, ad(b["YB"]), 1, 0) for b, _ in items] +
                    [(b["M"], 1, ad(b["idx"]), ad(ybar), 1, ad(b["YBb"]), 1, 1) for b, _ in items],
                    0)
        # tangent along V: zt_l = ht_{l-1} W_l, ht_l = s_l zt_l (rows M.. of H_l);
        # the last layer also starts the reverse pass,
        # zb_3 = s_3 ybar w_4 + a s_3 (1 - s_3) zt_3 w_4 (rows ..M of ZZ3)
        for l, (hin, hout, sk, zt) in enumerate(layers):
            grp = []
            for b, prm in items:
                M, ldh = b["M"], b["ldh"]
                ldp, Kp = (b["ldx"], D) if l == 0 else (ldh, H)
                g = dict(M=M, N=H, K=Kp, A=ad(b[hin], M, ldp), lda=ldp, B=ad(b["Wb"][l]), ldb=H,
                         C=ad(b[zt]), ldc=ldh, s=ad(b[sk]), aux=ad(b[hout], M, ldh))
                if l == 2:
                    g.update(C=None, x1=ad(prm[3][0]), x2=ad(b["YB"]), out2=ad(b["ZZ3"]))
                grp.append(g)
            self._ggemm(grp, 0, 0, 2, io=1 | 4 | 8)   # zt FP32; ht_l, zb_3 BF16
        # reverse: zb_{l-1} = s_{l-1} (zb_l W_l^T) + a s_{l-1} (1 - s_{l-1}) zt_{l-1} hd_{l-1}
        for (l, zz_in, sk, zt, hd, zz_out) in ((2, "ZZ3", "S2", "ZT2", "HD2", "ZZ2"),
                                               (1, "ZZ2", "S1", "ZT1", "HD1", "ZZ1")):
            self._ggemm([dict(M=b["M"], N=H, K=H, A=ad(b[zz_in]), lda=b["ldh"], B=ad(b["Wb"][l]),
                              ldb=H, C=ad(b[zz_out]), ldc=b["ldh"], s=ad(b[sk]), x1=ad(b[zt]),
                              x2=ad(b[hd])) for b, prm in items], 0, 1, 3, io=1 | 2)   # zb BF16
        # weight gradients over the stacked rows (bias from the ones column),
        # every layer of every element in one grouped launch
        grp = []
        for b, _ in items:
            M2, ldx, ldh = 2 * b["M"], b["ldx"], b["ldh"]
            gW1, gW2, gW3, gw4 = b["gW"]
            wk = b["work"]
            for (Mo, No, A, lda, B, ldb, G, w) in (
                    (D + 1, H, b["X0"], ldx, b["ZZ1"], ldh, gW1, wk[0]),
                    (H + 1, H, b["H1"], ldh, b["ZZ2"], ldh, gW2, wk[1]),
                    (H + 1, H, b["H2"], ldh, b["ZZ3"], ldh, gW3, wk[2]),
                    (H + 1, 1, b["H3"], ldh, b["YBb"], 1, gw4, wk[3])):
                grp.append(dict(M=Mo, N=No, K=M2, A=ad(A), lda=lda, B=ad(B), ldb=ldb, C=ad(G),
                                ldc=No, work=ad(w)))
        self._ggemm(grp, 1, 0, 0, nsplit=items[0][0]["nsplit"])
        grads = []
        for e in self.model.eles:
            b = bufs.get(e)
            if b is None:   # no atoms of this element in the batch
                grads += [torch.zeros_like(t) for (W, bb) in self.params[e] for t in (W, bb)]
                continue
            gW1, gW2, gW3, gw4 = b["gW"]
            grads += [gW1[:D], gW1[D], gW2[:H], gW2[H], gW3[:H], gW3[H], gw4[:H], gw4[H]]
        # parameters() order: element by element, (W, b) per layer; shapes as the parameters
        grads = [g.view(p.shape) for g, p in zip(grads, self.parameters())]
        if not views:
            grads = [g.clone() for g in grads]
        return L, E, F, grads

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
            L = self.step(tb, pos, check=False)
        return g, L
