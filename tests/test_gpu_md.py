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
