"""The tabulated Coulomb / vdW sweep (tm_coulomb_tab) against the analytic
sweep (tm_coulomb) and the FP64 oracle.

Systems on a cell grid: the periodic 1,000-water box (edge 31.06 A), the open
1,200-atom C/H/N/O cluster, and the 4,096-water box of the MD benchmark
(49.7 A).  Tolerances, from the tables' float64 fit
(|dK| < 3e-9, derivative 2e-5 of its maximum) and float32 lane sums:
  phi:   max |diff| < 2e-6 (|phi| ~ 0.1-1 per unit charge)
  grad:  rms |diff| < 2e-6 Ha/A, max < 2e-5
  eatom: summed Coulomb and vdW energies within 1e-7 Ha per atom
The oracle gates (per-atom energy 2e-6 Ha, force rms 1e-4 Ha/A) run through the
engine, which takes the tabulated sweep on these systems.
"""
import numpy as np
import pytest
import torch

from tests import helpers as H

pytestmark = pytest.mark.gpu


def _run(name, modelname, tables, dev):
    from tensormol_amd.engine import BPEEEngine
    if name == "box4096":
        from tensormol_amd import workloads as W
        Z, x, box = W.water_box(nside=16, edge=49.7, seed=1)
        ms, boxes, model = np.array([0, len(Z)]), box[None], H.model_for("water")
    else:
        Z, x, ms, boxes = H.system(name)
        model = H.model_for(modelname)
    eng = BPEEEngine(model, precision="fp32", coulomb_tables=tables)
    ws = eng.prepare(Z, ms, boxes, pos0=x)
    assert ws.use_cells
    pos = torch.tensor(x, dtype=torch.float64, device=dev)
    eng.run(ws, pos)
    torch.cuda.synchronize()
    return eng, ws, pos


@pytest.mark.parametrize("name,modelname", [("waterbox10", "water"), ("chno_small", "chno"),
                                            ("box4096", "water")])
def test_tab_sweep_matches_analytic_sweep(cuda, name, modelname):
    eng, ws, _ = _run(name, modelname, True, cuda)
    assert eng.coulomb_tables
    ref_eng, ref, _ = _run(name, modelname, False, cuda)
    N = ws.N
    dphi = (ws.phi[:N] - ref.phi[:N]).abs().max().item()
    dg = (ws.gradc[:3 * N] - ref.gradc[:3 * N]).view(N, 3)
    grms = dg.pow(2).mean().sqrt().item()
    gmax = dg.abs().max().item()
    de = (ws.eatom[:2 * N] - ref.eatom[:2 * N]).view(N, 2).sum(0).abs().cpu().numpy() / N
    print(f"{name}: |dphi| max {dphi:.2e} (|phi| max {ref.phi[:N].abs().max().item():.3f}), "
          f"grad diff rms {grms:.2e} max {gmax:.2e} Ha/A, energy diff per atom Ecc {de[0]:.2e} "
          f"Evdw {de[1]:.2e} Ha")
    assert dphi < 2e-6
    assert grms < 2e-6 and gmax < 2e-5
    assert np.all(de < 1e-7)


@pytest.mark.parametrize("name", ["waterbox10", "box4096"])
def test_tab_sweep_deterministic(cuda, name):
    """Two evaluations give the same bits: every sum is formed in an order
    fixed by the inputs."""
    eng, ws, pos = _run(name, "water", True, cuda)
    first = {k: getattr(ws, k).clone() for k in ("phi", "gradc", "eatom")}
    eng.run(ws, pos)
    torch.cuda.synchronize()
    for k, v in first.items():
        assert torch.equal(getattr(ws, k), v), k


@pytest.mark.parametrize("name,modelname", [("waterbox10", "water"), ("chno_small", "chno")])
def test_tab_sweep_vs_oracle(cuda, name, modelname):
    eng, ws, _ = _run(name, modelname, True, cuda)
    r = eng.results(ws)
    ref = H.oracle_eval(name, modelname)
    natom = np.diff(H.system(name)[2])
    dcc = np.abs(r["Ecc"] - ref["Ecc"]) / natom
    dvdw = np.abs(r["Evdw"] - ref["Evdw"]) / natom
    frms = np.sqrt(((r["gradient"] - ref["gradient"]) ** 2).mean())
    print(f"{name}: |dEcc| per atom {dcc.max():.2e}, |dEvdw| per atom {dvdw.max():.2e} Ha, "
          f"force rms err {frms:.2e} Ha/A")
    assert np.all(dcc < 2e-6) and np.all(dvdw < 1e-8)
    assert frms < 1e-4
