"""Velocity Verlet entry points through the C ABI (VelocityVerletStep,
TensorMol/Simulations/SimpleMD.py:14-38, split into drift and kick): the fused
kick-drift of one launch equals tm_vv_kick followed by tm_vv_drift, bit for
bit, on every output."""
import numpy as np
import pytest
import torch

from tensormol_amd import _lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N", [1, 1000, 12288])
def test_kick_drift_equals_kick_then_drift(cuda, N):
    L, P = _lib.lib(), _lib.ptr
    rng = np.random.default_rng(N)
    f = torch.tensor(rng.standard_normal((N, 3)) * 50.0, device=cuda)
    m = torch.tensor(rng.uniform(1.0, 16.0, N), device=cuda)
    v0 = torch.tensor(rng.standard_normal((N, 3)) * 1e-3, device=cuda)
    a0 = torch.tensor(rng.standard_normal((N, 3)) * 1e-9, device=cuda)
    x0 = torch.tensor(rng.uniform(0.0, 30.0, (N, 3)), device=cuda)
    dt = 0.5
    outs = []
    for fused in (False, True):
        v, a, x = v0.clone(), a0.clone(), x0.clone()
        ke = torch.zeros(N, dtype=torch.float64, device=cuda)
        if fused:
            assert L.tm_vv_kick_drift(N, dt, P(f), P(m), P(v), P(a), P(ke), P(x), None) == 0
        else:
            assert L.tm_vv_kick(N, dt, P(f), P(m), P(v), P(a), P(ke), None) == 0
            assert L.tm_vv_drift(N, dt, P(v), P(a), P(x), None) == 0
        torch.cuda.synchronize()
        outs.append((v, a, x, ke))
    for name, p, q in zip(("v", "a", "x", "ke"), *outs):
        assert torch.equal(p, q), name
    assert not torch.equal(outs[1][2], x0)   # the drift moved the atoms


def test_force_gather_vv_equals_gather_then_kick_drift(cuda):
    """BPEEEngine.adjoint(..., vv=) advances the Verlet state inside the force
    gather's launch (tm_force_gather_vv): forces, gradient, v, a, ke and the
    drifted positions equal tm_force_gather followed by tm_vv_kick_drift."""
    from tensormol_amd import params
    from tensormol_amd.engine import BPEEEngine
    from tests import helpers as H
    L, P = _lib.lib(), _lib.ptr
    Z, x, ms, boxes = H.system("waterbox10")
    eng = BPEEEngine(H.model_for("water"), precision="bf16")
    ws = eng.prepare(Z, ms, boxes, pos0=x)
    m = torch.tensor(params.atomic_masses(Z), dtype=torch.float64, device=cuda)
    rng = np.random.default_rng(0)
    v0 = torch.tensor(rng.standard_normal((ws.N, 3)) * 1e-3, device=cuda)
    a0 = torch.tensor(rng.standard_normal((ws.N, 3)) * 1e-9, device=cuda)
    dt = 0.5
    outs = []
    for fused in (False, True):
        pos = torch.tensor(x, dtype=torch.float64, device=cuda)
        v, a = v0.clone(), a0.clone()
        ke = torch.zeros(ws.N, dtype=torch.float64, device=cuda)
        eng.build_lists(ws, pos)
        eng.descriptors(ws, pos)
        eng.networks(ws)
        eng.electrostatics(ws, pos)
        if fused:
            eng.adjoint(ws, pos, vv=(dt, m, v, a, ke, pos))
        else:
            eng.adjoint(ws, pos)
            assert L.tm_vv_kick_drift(ws.N, dt, P(ws.force), P(m), P(v), P(a), P(ke), P(pos),
                                      eng._stream()) == 0
        torch.cuda.synchronize()
        outs.append([t.clone() for t in (ws.force, ws.grad, v, a, ke, pos)])
    for name, p, q in zip(("force", "grad", "v", "a", "ke", "x"), *outs):
        assert torch.equal(p, q), name
    assert not torch.equal(outs[1][5], torch.tensor(x, dtype=torch.float64, device=cuda))


def test_device_md_run_steps_equals_steps(cuda):
    """DeviceMD.run_steps(n) (kick and next drift inside the force gather) leaves
    positions, velocities, accelerations, kinetic energies and forces equal, bit
    for bit, to n calls of step() (drift, evaluation, kick)."""
    from tensormol_amd.engine import BPEEEngine
    from tensormol_amd.md import DeviceMD
    from tests import helpers as H
    Z, x, ms, boxes = H.system("waterbox10")
    eng = BPEEEngine(H.model_for("water"), precision="bf16")
    outs = []
    for fused in (False, True):
        ws = eng.prepare(Z, ms, boxes, pos0=x)
        md = DeviceMD(eng, ws, Z, x, dt=0.5, kind="nve", seed=3)
        md.start()
        if fused:
            md.run_steps(4)
        else:
            for _ in range(4):
                md.step()
        torch.cuda.synchronize()
        outs.append([t.clone() for t in (md.pos, md.v, md.a, md.ke, ws.force)])
    for name, p, q in zip(("x", "v", "a", "ke", "force"), *outs):
        assert torch.equal(p, q), name


# ------------------------------------------------------------------ oracle parity
def _gentle():
    """8 waters, 32-wide nets: a smooth potential over a short trajectory."""
    from oracle import oracle as O
    from tensormol_amd import workloads as W
    from tensormol_amd.model import BPEEModel
    Z, x = W.water_cluster(8, box=6.0, seed=1)
    model = BPEEModel([1, 8], dsf_alpha=0.2, seed=3, hidden=[32, 32, 32])
    sp = O.SymParams(model.eles)
    ms = np.array([0, len(Z)])

    def eandf(xx):
        r = O.evaluate(xx, Z, ms, sp, model.energy_nets, model.charge_nets, alpha=model.alpha,
                       dsf_alpha=model.dsf_alpha)
        return r["Etotal"][0], -O.JOULEPERHARTREE * r["gradient"]
    return Z, x, ms, model, eandf


@pytest.mark.parametrize("kind", ["nve", "nvt"])
def test_device_md_follows_oracle(cuda, kind):
    """DeviceMD on the FP32 path against the FP64 oracle's Velocity Verlet
    (SimpleMD.py:14-38) / Nose-Hoover step (SimpleMD.py:104-129), 24 steps of
    0.5 fs from the same state: positions within 1e-5 A, velocities within
    1e-4 relative, eta within 1e-3 relative, and (NVE) the total energy
    conserved as well as the oracle conserves it."""
    from oracle import oracle as O
    from tensormol_amd import params
    from tensormol_amd import workloads as W
    from tensormol_amd.engine import BPEEEngine
    from tensormol_amd.md import DeviceMD
    Z, x, ms, model, eandf = _gentle()
    m = params.atomic_masses(Z)
    v0 = W.maxwell_boltzmann(Z, 600.0, 5, m)
    eng = BPEEEngine(model, precision="fp32")
    ws = eng.prepare(Z, ms, None, pos0=x)
    md = DeviceMD(eng, ws, Z, x, v0=v0, dt=0.5, kind=kind, T=300.0)
    md.start()
    e0, f0 = eandf(x)
    xo, vo, ao, eta = x.copy(), v0.copy(), 1e-10 * f0 / m[:, None], 0.0
    et_gpu, et_orc = [], []
    nsteps = 24
    for k in range(nsteps):
        md.step()
        if kind == "nve":
            xo, vo, ao, eo = O.vv_step(xo, vo, ao, m, 0.5, eandf)
        else:
            xo, vo, ao, eo, eta = O.nose_step(xo, vo, ao, m, 0.5, eandf, eta, 300.0, 0.5)
        pe, ke = md.energies()
        et_gpu.append(pe[0] + ke[0])
        et_orc.append(eo + O.kinetic_energy(vo, m) * len(m) / O.JOULEPERHARTREE)
    torch.cuda.synchronize()
    xg, vg = md.pos.cpu().numpy(), md.v.cpu().numpy()
    dx = np.abs(xg - xo).max()
    dv = np.abs(vg - vo).max() / np.abs(vo).max()
    print(f"{kind}: |dx| {dx:.2e} A, |dv|/|v| {dv:.2e}")
    assert dx < 1e-5 and dv < 1e-4
    if kind == "nvt":
        eg = float(md.eta.cpu()[0])
        print(f"eta gpu {eg:.6e} oracle {eta:.6e}")
        assert eta != 0.0 and abs(eg - eta) < 1e-3 * abs(eta)
    else:
        et_gpu, et_orc = np.array(et_gpu), np.array(et_orc)
        print(f"Etot spread gpu {np.ptp(et_gpu):.2e} oracle {np.ptp(et_orc):.2e} Ha")
        assert np.abs(et_gpu - et_orc).max() < 1e-4
        assert np.ptp(et_gpu) < 1.5 * np.ptp(et_orc) + 1e-5


def test_device_md_raises_on_row_overflow(cuda):
    """A neighbour stride far too small for the system: the rows are clamped
    by the kernels and flagged; DeviceMD.step() raises instead of carrying on
    with truncated rows."""
    from tensormol_amd.engine import BPEEEngine, NeighbourOverflow
    from tensormol_amd.md import DeviceMD
    Z, x, ms, model, _ = _gentle()
    eng = BPEEEngine(model, precision="fp32", ldr=4, lda=4, check_overflow=False)
    ws = eng.prepare(Z, ms, None, pos0=x)
    md = DeviceMD(eng, ws, Z, x, dt=0.5, kind="nve")
    md.start()
    with pytest.raises(NeighbourOverflow):
        md.step()
    md2 = DeviceMD(eng, ws, Z, x, dt=0.5, kind="nve")
    md2.start()
    md2.run_steps(2, check=False)   # no check asked: no raise
    with pytest.raises(NeighbourOverflow):
        md2.check_flags()
