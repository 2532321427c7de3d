"""Training path on the GPU: the descriptor tangent kernel against the FP64
oracle, the adjoint identity <A u, J> = <u, A^T J> between the tangent and
adjoint kernels, and the force-matching weight gradient against the FP64
oracle's double-backward gradient and against central finite differences of
the oracle loss on 32 random weights.

Tolerances: tangent 2e-5 relative to its largest entry (FP32 descriptors);
adjoint identity 1e-5 relative; FP32 training gradient 1e-4 relative (vector
norm) and per weight |g - fd| <= 1e-5 max|g|; BF16-operand forward 3e-2.
"""
import functools

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tensormol_amd import workloads as W
from tensormol_amd.model import BPEEModel

pytestmark = pytest.mark.gpu

ELES = [1, 6, 7, 8]


@functools.lru_cache(maxsize=None)
def _setup():
    mols = W.chain_set(np.array([20, 27, 33, 41]), seed=9)
    Z, x, ms = W.pack(mols)
    sp = O.SymParams(ELES)
    lists = O.build_lists(x, Z, ms, sp)
    teacher = BPEEModel(ELES, dsf_alpha=0.2, seed=1, hidden=[64, 64, 64])
    student = BPEEModel(ELES, dsf_alpha=0.2, seed=2, hidden=[64, 64, 64])
    xt = torch.tensor(x, requires_grad=True)
    Et = O.bp_energies(xt, Z, ms, sp, O.layers_torch(teacher.energy_nets), teacher.alpha, lists)
    (g,) = torch.autograd.grad(Et.sum(), xt)
    B, N = len(ms) - 1, len(Z)
    w_e, w_f = np.full(B, 1.0 / B), np.full(B, 0.5 / (3 * N))
    return Z, x, ms, sp, lists, student, Et.detach().numpy(), -g.numpy(), w_e, w_f


def _oracle_grad(student, Z, x, ms, sp, lists, E_ref, F_ref, w_e, w_f):
    lt = O.layers_torch(student.energy_nets, requires_grad=True)
    L, _, _ = O.force_matching_loss(x, Z, ms, sp, lt, student.alpha, E_ref, F_ref, w_e, w_f, lists)
    params = [t for e in student.eles for (Wt, bt) in lt[e] for t in (Wt, bt)]
    return float(L), [g.numpy() for g in torch.autograd.grad(L, params)]


def test_descriptor_jvp_vs_oracle(cuda):
    from tensormol_amd.engine import BPEEEngine
    Z, x, ms, sp, lists, student, *_ = _setup()
    eng = BPEEEngine(student, precision="fp32")
    ws = eng.prepare(Z, ms, None)
    pos = torch.tensor(x, dtype=torch.float64, device=cuda)
    eng.build_lists(ws, pos)
    eng.descriptors(ws, pos)
    du = np.random.default_rng(1).normal(size=x.shape)
    V = torch.zeros(ws.N, eng.D, dtype=torch.float32, device=cuda)
    eng.descriptor_jvp(ws, pos, torch.tensor(du, device=cuda), V)
    ref = O.descriptor_jvp(x, Z, sp, du, lists)
    err = np.abs(V.cpu().numpy() - ref).max() / np.abs(ref).max()
    print(f"tangent max err {err:.2e} (relative to max |V| {np.abs(ref).max():.3f})")
    assert err < 2e-5


def test_adjoint_identity(cuda):
    from tensormol_amd.engine import BPEEEngine
    Z, x, ms, sp, lists, student, *_ = _setup()
    eng = BPEEEngine(student, precision="fp32")
    ws = eng.prepare(Z, ms, None)
    pos = torch.tensor(x, dtype=torch.float64, device=cuda)
    eng.build_lists(ws, pos)
    rng = np.random.default_rng(2)
    u = torch.tensor(rng.normal(size=x.shape), device=cuda)
    J = torch.tensor(rng.normal(size=(ws.N, eng.D)), dtype=torch.float32, device=cuda)
    V = torch.zeros(ws.N, eng.D, dtype=torch.float32, device=cuda)
    g = torch.zeros(ws.N, 3, dtype=torch.float64, device=cuda)
    eng.descriptor_jvp(ws, pos, u, V)
    eng.descriptor_vjp(ws, pos, J, g)
    lhs = float((V.double() * J.double()).sum())
    rhs = float((u * g).sum())
    print(f"<Au, J> = {lhs:.8e}, <u, A^T J> = {rhs:.8e}")
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), abs(rhs))


@pytest.mark.parametrize("dtype,tol", [("fp32", 1e-4), ("bf16", 3e-2)])
def test_training_gradient_vs_oracle(cuda, dtype, tol):
    from tensormol_amd.train import ForceMatchingTrainer
    Z, x, ms, sp, lists, student, E_ref, F_ref, w_e, w_f = _setup()
    tr = ForceMatchingTrainer(student, mlp_dtype=dtype)
    tb = tr.batch(Z, ms, E_ref, F_ref, w_e, w_f)
    pos = torch.tensor(x, dtype=torch.float64, device=cuda)
    L, grads = tr.gradients(tb, pos)
    L_ref, g_ref = _oracle_grad(student, Z, x, ms, sp, lists, E_ref, F_ref, w_e, w_f)
    gv = np.concatenate([g.double().cpu().numpy().ravel() for g in grads])
    rv = np.concatenate([g.ravel() for g in g_ref])
    err = np.linalg.norm(gv - rv) / np.linalg.norm(rv)
    print(f"{dtype}: loss {float(L):.6e} vs oracle {L_ref:.6e}; gradient rel err {err:.2e}")
    assert abs(float(L) - L_ref) <= tol * abs(L_ref)
    assert err < tol


def test_mfma_gradient_vs_torch_bf16(cuda):
    """The hand-written MFMA path against torch on the same BF16-rounded forward
    operands (the MFMA path also rounds the backward GEMMs' gradient operands):
    the two weight gradients agree to 1e-2 of their norm, losses to 1e-4."""
    from tensormol_amd.train import ForceMatchingTrainer
    Z, x, ms, sp, lists, student, E_ref, F_ref, w_e, w_f = _setup()
    pos = torch.tensor(x, dtype=torch.float64, device=cuda)
    out = {}
    for dt in ("bf16", "bf16_torch"):
        tr = ForceMatchingTrainer(student, mlp_dtype=dt)
        tb = tr.batch(Z, ms, E_ref, F_ref, w_e, w_f)
        L, grads = tr.gradients(tb, pos)
        out[dt] = (float(L), np.concatenate([g.double().cpu().numpy().ravel() for g in grads]))
    (La, ga), (Lb, gb) = out["bf16"], out["bf16_torch"]
    err = np.linalg.norm(ga - gb) / np.linalg.norm(gb)
    print(f"MFMA vs torch bf16: loss {La:.6e} / {Lb:.6e}, gradient rel diff {err:.2e}")
    assert abs(La - Lb) <= 1e-4 * abs(Lb)
    assert err < 1e-2


def test_training_gradient_finite_differences(cuda):
    """32 random weights: GPU gradient vs central differences of the FP64 loss."""
    from tensormol_amd.train import ForceMatchingTrainer
    Z, x, ms, sp, lists, student, E_ref, F_ref, w_e, w_f = _setup()
    tr = ForceMatchingTrainer(student)
    tb = tr.batch(Z, ms, E_ref, F_ref, w_e, w_f)
    _, grads = tr.gradients(tb, torch.tensor(x, dtype=torch.float64, device=cuda))
    grads = [g.double().cpu().numpy() for g in grads]
    gmax = max(np.abs(g).max() for g in grads)
    rng = np.random.default_rng(5)
    h = 1e-5
    worst = 0.0
    for _ in range(32):
        k = int(rng.integers(len(grads)))
        idx = tuple(int(rng.integers(n)) for n in grads[k].shape)
        vals = []
        for sgn in (1.0, -1.0):
            lp = O.layers_torch(student.energy_nets)
            flat = [t for e in student.eles for (Wt, bt) in lp[e] for t in (Wt, bt)]
            with torch.no_grad():
                flat[k][idx] += sgn * h
            Lp, _, _ = O.force_matching_loss(x, Z, ms, sp, lp, student.alpha, E_ref, F_ref, w_e,
                                             w_f, lists)
            vals.append(float(Lp.detach()))
        fd = (vals[0] - vals[1]) / (2 * h)
        worst = max(worst, abs(grads[k][idx] - fd) / gmax)
    print(f"32 weights: max |g - fd| / max|g| = {worst:.2e}")
    assert worst < 1e-5


def test_training_steps_reduce_loss(cuda):
    from tensormol_amd.train import ForceMatchingTrainer
    Z, x, ms, sp, lists, student, E_ref, F_ref, w_e, w_f = _setup()
    tr = ForceMatchingTrainer(student, lr=1e-3)
    tb = tr.batch(Z, ms, E_ref, F_ref, w_e, w_f)
    pos = torch.tensor(x, dtype=torch.float64, device=cuda)
    losses = [float(tr.step(tb, pos)) for _ in range(20)]
    print("losses", losses[0], losses[-1])
    assert losses[-1] < 0.5 * losses[0]


@pytest.mark.parametrize("dtype", ["fp32", "bf16"])
def test_captured_step_matches_eager(cuda, dtype):
    """A training step captured into a HIP graph and replayed six times follows
    the eager steps from the same state: every replay's loss and the final
    weights agree to 1e-5 relative (index_add's atomic order is the only
    difference), which also rules out graph nodes whose replays go wrong."""
    from tensormol_amd.train import ForceMatchingTrainer
    Z, x, ms, sp, lists, student, E_ref, F_ref, w_e, w_f = _setup()
    pos = torch.tensor(x, dtype=torch.float64, device=cuda)
    eager = ForceMatchingTrainer(student, lr=1e-3, mlp_dtype=dtype, capturable=True)
    graphed = ForceMatchingTrainer(student, lr=1e-3, mlp_dtype=dtype, capturable=True)
    tb_e = eager.batch(Z, ms, E_ref, F_ref, w_e, w_f)
    tb_g = graphed.batch(Z, ms, E_ref, F_ref, w_e, w_f)
    g, L = graphed.capture(tb_g, pos, warmup=3)
    for _ in range(3):   # the capture's warm-up steps
        eager.step(tb_e, pos)
    for k in range(6):
        le = float(eager.step(tb_e, pos))
        g.replay()
        torch.cuda.synchronize()
        lg = float(L)
        print(f"step {k}: eager {le:.8e} graph {lg:.8e}")
        assert abs(lg - le) <= 1e-5 * abs(le), (k, le, lg)
    for a, b in zip(eager.parameters(), graphed.parameters()):
        assert torch.allclose(a, b, rtol=1e-5, atol=1e-7)


# ---------------------------------------------------------------- config [4] shape
@functools.lru_cache(maxsize=None)
def _setup_config4():
    """The force-matching workload at its stated shape: 768-512-512-512-1
    teacher and student nets (C/H/N/O, softplus), 16 chain molecules of
    20-60 atoms built as in the serving stream (seed 4)."""
    sizes = np.random.default_rng(4).integers(20, 61, size=16)
    Z, x, ms = W.pack(W.chain_set(sizes, seed=4))
    sp = O.SymParams(ELES)
    lists = O.build_lists(x, Z, ms, sp)
    teacher = BPEEModel(ELES, dsf_alpha=0.2, seed=1, hidden=[512, 512, 512])
    student = BPEEModel(ELES, dsf_alpha=0.2, seed=2, hidden=[512, 512, 512])
    assert student.sym.D == 768
    xt = torch.tensor(x, requires_grad=True)
    Et = O.bp_energies(xt, Z, ms, sp, O.layers_torch(teacher.energy_nets), teacher.alpha, lists)
    (g,) = torch.autograd.grad(Et.sum(), xt)
    B, N = len(ms) - 1, len(Z)
    w_e, w_f = np.full(B, 1.0 / B), np.full(B, 0.5 / (3 * N))
    return Z, x, ms, sp, lists, student, Et.detach().numpy(), -g.numpy(), w_e, w_f


@pytest.mark.parametrize("dtype,tol", [("fp32", 2e-4), ("bf16", 3e-2)])
def test_config4_gradient_vs_oracle(cuda, dtype, tol):
    from tensormol_amd.train import ForceMatchingTrainer
    Z, x, ms, sp, lists, student, E_ref, F_ref, w_e, w_f = _setup_config4()
    tr = ForceMatchingTrainer(student, mlp_dtype=dtype)
    tb = tr.batch(Z, ms, E_ref, F_ref, w_e, w_f)
    L, grads = tr.gradients(tb, torch.tensor(x, dtype=torch.float64, device=cuda))
    L_ref, g_ref = _oracle_grad(student, Z, x, ms, sp, lists, E_ref, F_ref, w_e, w_f)
    gv = np.concatenate([g.double().cpu().numpy().ravel() for g in grads])
    rv = np.concatenate([g.ravel() for g in g_ref])
    err = np.linalg.norm(gv - rv) / np.linalg.norm(rv)
    print(f"config [4] {dtype}: {len(Z)} atoms, loss {float(L):.6e} vs oracle {L_ref:.6e}; "
          f"gradient rel err {err:.2e}")
    assert abs(float(L) - L_ref) <= tol * abs(L_ref)
    assert err < tol


def test_config4_gradient_finite_differences(cuda):
    """32 random weights of the 768-512-512-512-1 student: GPU gradient vs
    central differences of the FP64 oracle loss."""
    from tensormol_amd.train import ForceMatchingTrainer
    Z, x, ms, sp, lists, student, E_ref, F_ref, w_e, w_f = _setup_config4()
    tr = ForceMatchingTrainer(student)
    tb = tr.batch(Z, ms, E_ref, F_ref, w_e, w_f)
    _, grads = tr.gradients(tb, torch.tensor(x, dtype=torch.float64, device=cuda))
    grads = [g.double().cpu().numpy() for g in grads]
    gmax = max(np.abs(g).max() for g in grads)
    rng = np.random.default_rng(6)
    h = 1e-5
    worst = 0.0
    for _ in range(32):
        k = int(rng.integers(len(grads)))
        idx = tuple(int(rng.integers(n)) for n in grads[k].shape)
        vals = []
        for sgn in (1.0, -1.0):
            lp = O.layers_torch(student.energy_nets)
            flat = [t for e in student.eles for (Wt, bt) in lp[e] for t in (Wt, bt)]
            with torch.no_grad():
                flat[k][idx] += sgn * h
            Lp, _, _ = O.force_matching_loss(x, Z, ms, sp, lp, student.alpha, E_ref, F_ref, w_e,
                                             w_f, lists)
            vals.append(float(Lp.detach()))
        fd = (vals[0] - vals[1]) / (2 * h)
        worst = max(worst, abs(grads[k][idx] - fd) / gmax)
    print(f"config [4], 32 weights: max |g - fd| / max|g| = {worst:.2e}")
    assert worst < 1e-4
