"""Pins the oracle's training pieces: the descriptor tangent against central
finite differences of the descriptors, and the force-matching loss gradient in
the weights against central finite differences of the loss (both float64)."""
import numpy as np
import torch

from oracle import oracle as O
from tensormol_amd import workloads as W
from tensormol_amd.model import BPEEModel


def _batch(sizes=(9, 12), seed=4):
    mols = W.chain_set(np.array(sizes), seed=seed)
    return W.pack(mols)


def _model(seed, hidden=(16, 16, 16)):
    return BPEEModel([1, 6, 7, 8], dsf_alpha=0.2, seed=seed, hidden=list(hidden))


def test_descriptor_jvp_central_differences():
    Z, x, ms = _batch()
    sp = O.SymParams([1, 6, 7, 8])
    lists = O.build_lists(x, Z, ms, sp)
    du = np.random.default_rng(0).normal(size=x.shape)
    jv = O.descriptor_jvp(x, Z, sp, du, lists)
    h = 1e-5
    g = lambda y: O.descriptors(torch.tensor(y), Z, sp, lists).numpy()
    fd = (g(x + h * du) - g(x - h * du)) / (2 * h)
    err = np.abs(jv - fd).max() / np.abs(fd).max()
    assert err < 1e-6, err
    assert np.abs(jv).max() > 1e-3     # a non-trivial tangent


def test_force_matching_gradient_central_differences():
    Z, x, ms = _batch()
    sp = O.SymParams([1, 6, 7, 8])
    lists = O.build_lists(x, Z, ms, sp)
    teacher, student = _model(1), _model(2)
    B, N = len(ms) - 1, len(Z)
    xt = torch.tensor(x, requires_grad=True)
    Et = O.bp_energies(xt, Z, ms, sp, O.layers_torch(teacher.energy_nets), teacher.alpha, lists)
    (gt,) = torch.autograd.grad(Et.sum(), xt)
    E_ref, F_ref = Et.detach().numpy(), -gt.numpy()
    w_e, w_f = np.full(B, 1.0 / B), np.full(B, 0.5 / (3 * N))
    lt = O.layers_torch(student.energy_nets, requires_grad=True)
    L, _, _ = O.force_matching_loss(x, Z, ms, sp, lt, student.alpha, E_ref, F_ref, w_e, w_f, lists)
    params = [t for e in sorted(lt) for (Wt, bt) in lt[e] for t in (Wt, bt)]
    grads = torch.autograd.grad(L, params)
    rng = np.random.default_rng(3)
    h = 1e-5
    for _ in range(8):
        k = int(rng.integers(len(params)))
        idx = tuple(int(rng.integers(n)) for n in params[k].shape)
        vals = []
        for sgn in (1.0, -1.0):
            lp = O.layers_torch(student.energy_nets)
            flat = [t for e in sorted(lp) for (Wt, bt) in lp[e] for t in (Wt, bt)]
            with torch.no_grad():
                flat[k][idx] += sgn * h
            Lp, _, _ = O.force_matching_loss(x, Z, ms, sp, lp, student.alpha, E_ref, F_ref, w_e,
                                             w_f, lists)
            vals.append(float(Lp))
        fd = (vals[0] - vals[1]) / (2 * h)
        an = float(grads[k][idx])
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)) + 1e-9, (k, idx, fd, an)
    # the teacher's own weights give zero loss and zero gradient
    lt0 = O.layers_torch(teacher.energy_nets, requires_grad=True)
    L0, _, _ = O.force_matching_loss(x, Z, ms, sp, lt0, teacher.alpha, E_ref, F_ref, w_e, w_f,
                                     lists)
    assert float(L0) < 1e-20
