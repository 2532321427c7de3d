"""Fused MFMA MLP kernel vs a plain PyTorch FP32 reference of the same op, and
end-to-end reduced-precision E+F vs the FP64 oracle.

Tolerances (relative to the RMS of the reference quantity):
  bf16 (8-bit mantissa operands): y 2e-2, dy/dG 3e-2
  fp16 (11-bit):                  y 3e-3, dy/dG 5e-3
  bf16x3 (hi/lo split):           y 2e-4, dy/dG 5e-4
End-to-end gate of the benchmark: force RMS error < 1e-3 Ha/A vs the FP64 oracle.
"""
import numpy as np
import pytest
import torch

from tests import helpers as H

pytestmark = pytest.mark.gpu

TOL = {"bf16": (2e-2, 3e-2), "fp16": (3e-3, 5e-3), "bf16x3": (2e-4, 5e-4)}


def torch_mlp_ref(G, layers, alpha):
    """Plain FP32 reference: y = MLP(G), J = dy/dG via autograd."""
    x = G.clone().requires_grad_(True)
    h = x
    for (W, b) in layers[:-1]:
        z = h @ torch.tensor(W, dtype=torch.float32, device=G.device) + torch.tensor(
            b, dtype=torch.float32, device=G.device)
        h = torch.nn.functional.softplus(z * alpha) / alpha
    W, b = layers[-1]
    y = (h @ torch.tensor(W, dtype=torch.float32, device=G.device))[:, 0] + float(b[0])
    (J,) = torch.autograd.grad(y.sum(), x)
    return y.detach(), J


@pytest.mark.parametrize("precision", ["bf16", "fp16", "bf16x3"])
@pytest.mark.parametrize("name,modelname", [("water64", "water"), ("chains", "chno")])
def test_mfma_mlp_vs_torch_fp32(cuda, precision, name, modelname):
    from tensormol_amd.engine import BPEEEngine
    model = H.model_for(modelname)
    eng = BPEEEngine(model, precision=precision)
    Z, x, ms, boxes = H.system(name)
    ws = eng.prepare(Z, ms, boxes, pos0=x)
    pos = torch.tensor(x, dtype=torch.float64, device=cuda)
    eng.build_lists(ws, pos)
    eng.descriptors(ws, pos)
    eng.networks(ws)
    torch.cuda.synchronize()
    ty, tj = TOL[precision]
    for net, nets in ((0, model.energy_nets), (1, model.charge_nets)):
        y_ref = torch.zeros(ws.N, device=cuda)
        J_ref = torch.zeros(ws.N, eng.D, device=cuda)
        for a, e in enumerate(model.eles):
            idx = torch.tensor(np.nonzero(ws.elem_np == a)[0], device=cuda)
            if len(idx) == 0:
                continue
            y, J = torch_mlp_ref(ws.G[idx], nets[e], model.alpha)
            y_ref[idx] = y
            J_ref[idx] = J
        y_gpu = ws.eps if net == 0 else ws.qraw
        J_gpu = ws.JE if net == 0 else ws.JQ
        ey = ((y_gpu - y_ref).pow(2).mean().sqrt() / y_ref.pow(2).mean().sqrt()).item()
        ej = ((J_gpu - J_ref).pow(2).mean().sqrt() / J_ref.pow(2).mean().sqrt()).item()
        print(f"{precision} {name} net{net}: rel rms err y {ey:.2e} dy/dG {ej:.2e}")
        assert ey < ty and ej < tj


@pytest.mark.parametrize("precision", ["bf16", "fp16", "bf16x3"])
@pytest.mark.parametrize("name,modelname", [("water64", "water"), ("chains", "chno"),
                                            ("waterbox10", "water"), ("chno_small", "chno"),
                                            ("johnsonmols", "chno"), ("peptide188", "chno")])
def test_energy_forces_mfma(cuda, precision, name, modelname):
    from tensormol_amd.engine import BPEEEngine
    eng = BPEEEngine(H.model_for(modelname), precision=precision)
    Z, x, ms, boxes = H.system(name)
    ws = eng.prepare(Z, ms, boxes, pos0=x)
    pos = torch.tensor(x, dtype=torch.float64, device=cuda)
    eng.run(ws, pos)
    torch.cuda.synchronize()
    r = eng.results(ws)
    ref = H.oracle_eval(name, modelname)
    natom = np.diff(ms)
    dE = np.abs(r["Etotal"] - ref["Etotal"])
    frms = np.sqrt(((r["gradient"] - ref["gradient"]) ** 2).mean())
    print(f"{precision} {name}: |dE|/atom max {np.max(dE / natom):.2e} Ha, |dE| max {dE.max():.2e} Ha, "
          f"force rms err {frms:.2e} Ha/A")
    assert frms < 1e-3
    if precision == "bf16x3":
        assert np.all(dE / natom < 2e-5)


@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_tile_split_bit_identical(cuda, precision):
    """Whole rounds of 64-row tiles plus 32-row tiles for the rest give the
    same outputs, bit for bit, as 32-row tiles only: every output row's
    arithmetic is independent of the tile it sits in."""
    from tensormol_amd.engine import BPEEEngine
    Z, x, ms, boxes = H.system("waterbox10")
    pos = torch.tensor(x, dtype=torch.float64, device=cuda)
    outs = []
    for ncu in (1 << 20, 40):   # all 32-row tiles / two 64-row rounds of 40 tiles + 32-row rest
        eng = BPEEEngine(H.model_for("water"), precision=precision)
        eng.mfma.ncu = ncu
        ws = eng.prepare(Z, ms, boxes, pos0=x)
        if ncu == 40:
            assert ws.mlp_ws["n64"] > 0 and ws.mlp_ws["n32"] > 0
        eng.build_lists(ws, pos)
        eng.descriptors(ws, pos)
        eng.networks(ws)
        torch.cuda.synchronize()
        outs.append((ws.Y2[:2 * ws.N].clone(), ws.J2[:2 * ws.N * eng.D].clone()))
    assert torch.equal(outs[0][0], outs[1][0])
    assert torch.equal(outs[0][1], outs[1][1])
