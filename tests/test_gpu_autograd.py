"""The autograd wrapper: gradient = adjoint-kernel dE/dx, checked by central
finite differences of the engine's own energy and against the FP64 oracle.

Finite differences: h = 5e-3 A on the plain FP32 path.  The energy noise of
one perturbed coordinate is ~2e-7 Ha (FP32 per-atom terms of ~30 atoms summed
in FP64), i.e. ~2e-5 Ha/A after dividing by 2h, and the truncation error
h^2/6 |d3E/dx3| is ~1e-4 Ha/A for third derivatives up to 25 Ha/A^3; the
gate is 1e-3 Ha/A, the force tolerance of the engine.
"""
import numpy as np
import pytest
import torch

from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,modelname", [("water64", "water"), ("chains", "chno")])
def test_autograd_gradient_vs_central_differences(cuda, name, modelname):
    from tensormol_amd.autograd import energy
    from tensormol_amd.engine import BPEEEngine
    eng = BPEEEngine(H.model_for(modelname), precision="fp32")
    Z, x, ms, boxes = H.system(name)
    ws = eng.prepare(Z, ms, boxes, pos0=x)
    pos = torch.tensor(x, dtype=torch.float64, device=cuda, requires_grad=True)
    E = energy(eng, ws, pos)
    w = torch.linspace(0.5, 1.5, ws.B, dtype=torch.float64, device=cuda)
    (g,) = torch.autograd.grad((w * E).sum(), pos)
    g = g.cpu().numpy()
    wn = w.cpu().numpy()
    mol = np.repeat(np.arange(ws.B), np.diff(ms))
    rng = np.random.default_rng(0)
    picks = rng.choice(len(Z) * 3, size=12, replace=False)
    h = 5e-3
    errs = []
    for c in picks:
        a, d = divmod(int(c), 3)
        vals = []
        for sgn in (1.0, -1.0):
            xp = x.copy()
            xp[a, d] += sgn * h
            with torch.no_grad():
                Ep = energy(eng, ws, torch.tensor(xp, dtype=torch.float64, device=cuda))
            vals.append(float((torch.tensor(wn, device=cuda) * Ep).sum()))
        fd = (vals[0] - vals[1]) / (2 * h)
        errs.append(abs(fd - g[a, d]))
        print(f"{name} atom {a} dim {d} (molecule {mol[a]}): fd {fd:+.6e} autograd {g[a, d]:+.6e}")
    print(f"{name}: max |fd - autograd| {max(errs):.2e} Ha/A")
    assert max(errs) < 1e-3


def test_autograd_matches_oracle_gradient(cuda):
    from tensormol_amd.autograd import energy
    from tensormol_amd.engine import BPEEEngine
    eng = BPEEEngine(H.model_for("water"), precision="fp32")
    Z, x, ms, boxes = H.system("water64")
    ws = eng.prepare(Z, ms, boxes, pos0=x)
    pos = torch.tensor(x, dtype=torch.float64, device=cuda, requires_grad=True)
    E = energy(eng, ws, pos)
    (g,) = torch.autograd.grad(E.sum(), pos)
    ref = H.oracle_eval("water64", "water")
    err = np.sqrt(((g.cpu().numpy() - ref["gradient"]) ** 2).mean())
    print(f"autograd vs oracle gradient rms err {err:.2e} Ha/A")
    assert err < 1e-5
    assert abs(float(E[0]) - ref["Etotal"][0]) < 1e-4


def test_autograd_rejects_bad_input(cuda):
    from tensormol_amd.autograd import energy
    from tensormol_amd.engine import BPEEEngine
    eng = BPEEEngine(H.model_for("water"), precision="fp32")
    Z, x, ms, boxes = H.system("water64")
    ws = eng.prepare(Z, ms, boxes, pos0=x)
    with pytest.raises(TypeError):
        energy(eng, ws, torch.tensor(x, dtype=torch.float32, device=cuda))
    with pytest.raises(ValueError):
        energy(eng, ws, torch.tensor(x[:-1], dtype=torch.float64, device=cuda))
