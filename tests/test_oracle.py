"""Pins of the FP64 oracle (CPU only).

The reference ships no tests or golden vectors for this path and TensorFlow is
absent, so each oracle function is pinned by checks chosen so that a dropped
term, a wrong sign or index or a transposed operand fails: brute-force set
definitions, closed forms at special geometries, finite differences,
invariances and conservation laws.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tensormol_amd import workloads as W
from tensormol_amd.engine import nl_threshold


def _ulp_steps(x, k):
    return np.nextafter(x, np.inf) if k > 0 else np.nextafter(x, -np.inf)


# ------------------------------------------------------------ neighbour sets
@pytest.mark.parametrize("rc", [4.6, 3.1, 15.0])
def test_threshold_equals_reference_predicate(rc):
    thr = nl_threshold(rc)
    pred = lambda s: math.sqrt(s) + 0.0000000000001 < rc
    assert not pred(thr) and pred(float(np.nextafter(thr, 0.0)))
    s = thr
    for _ in range(200):
        s = float(np.nextafter(s, 0.0))
    for _ in range(400):
        assert pred(s) == (s < thr)
        s = float(np.nextafter(s, np.inf))


def _python_pairs(x, rc):
    n = len(x)
    out = [[] for _ in range(n)]
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            dx, dy, dz = x[i] - x[j]
            d = math.sqrt((dx * dx + dy * dy) + dz * dz) + 0.0000000000001
            if d < rc:
                out[i].append(j)
    return out


def test_nlist_matches_python_double_loop():
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 8, size=(60, 3))
    ref = _python_pairs(x, 3.1)
    got = O.nlist_naive(x, 3.1)
    for i in range(60):
        assert list(got[i]) == ref[i]


def test_nlist_ulp_adversarial():
    """Atoms at distances straddling the cutoff by single ulps."""
    rc = 4.6
    base = rc - 0.0000000000001
    ds = [base]
    for k in (1, 2, 3):
        ds.append(float(np.nextafter(ds[-1], np.inf)))
    lo = [base]
    for k in (1, 2, 3):
        lo.append(float(np.nextafter(lo[-1], -np.inf)))
    dist = sorted(set(ds + lo))
    x = np.zeros((len(dist) + 1, 3))
    for q, d in enumerate(dist):
        x[q + 1, 0] = 1000.0 * (q + 1)   # far apart from each other
    for q, d in enumerate(dist):
        xx = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0]])
        got = O.nlist_naive(xx, rc)
        expect = math.sqrt(d * d) + 0.0000000000001 < rc
        assert (len(got[0]) == 1) == expect


def test_full_list_symmetric_and_half_list():
    Z, x = W.water_cluster(20, seed=4)
    full = O.nlist_naive(x, 4.6, doperms=1)
    half = O.nlist_naive(x, 4.6, doperms=0)
    for i, row in enumerate(full):
        for j in row:
            assert i in full[j]
    assert sum(map(len, full)) == 2 * sum(map(len, half))
    for i, row in enumerate(half):
        assert np.all(row > i)


def test_periodic_min_image_matches_explicit_images():
    """Minimum-image set == pairs found among the 27 explicit images."""
    rng = np.random.default_rng(5)
    L = 10.0
    x = rng.uniform(0, L, size=(40, 3))
    got = O.nlist_naive(x, 4.6, box=np.array([L, L, L]))
    shifts = np.array([[a, b, c] for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]) * L
    for i in range(40):
        ref = set()
        for j in range(40):
            if j == i:
                continue
            d = np.sqrt((((x[i] - x[j])[None, :] - shifts) ** 2).sum(1)).min()
            if d + 1e-13 < 4.6:
                ref.add(j)
        assert set(got[i].tolist()) == ref


def test_pairs_triples_hand_built():
    """Centre O(0) with H(1), H(2), O(3) inside 3.1 A: the reference stores the
    lower-Z atom first and sorts by (m, i, l, k, j) (Neighbors.py:170-186, 380-383)."""
    x = np.array([[0.0, 0, 0], [1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.2]])
    Z = np.array([8, 1, 1, 8])
    trp, trt, mil, jkmax = O.pairs_triples_with_ele(x[None], np.array([4]), Z[None], [1, 8], 4.6, 3.1)
    t0 = trt[trt[:, 1] == 0]
    np.testing.assert_array_equal(t0, [[0, 0, 1, 2, 0], [0, 0, 1, 3, 1], [0, 0, 2, 3, 1]])
    np.testing.assert_array_equal(mil[trt[:, 1] == 0][:, 3], [0, 0, 1])
    p0 = trp[trp[:, 1] == 0]
    np.testing.assert_array_equal(p0, [[0, 0, 1, 0], [0, 0, 2, 0], [0, 0, 3, 1]])
    # centre H(1): neighbours O(0) d=1, H(2) d=sqrt2, O(3) d=1.56 -> pair (H2,O0)
    # stored H first
    t1 = trt[trt[:, 1] == 1]
    assert [2, 0] in t1[:, 2:4].tolist()


# --------------------------------------------------------------- descriptors
def test_radial_two_atom_closed_form():
    sp = O.SymParams([1, 8])
    r = 1.7
    x = torch.tensor([[0.0, 0, 0], [r, 0, 0]], dtype=torch.float64)
    G = O.sym_radial(x, np.array([0, 1]), np.array([1, 0]), np.array([1, 0]), 2, sp).numpy()
    fc = 0.5 * (math.cos(3.14159265359 * r / 4.6) + 1.0)
    expect = np.array([math.exp(-4.0 * (r - 4.6 * s / 32) ** 2) * fc for s in range(32)])
    np.testing.assert_allclose(G[0, 32:64], expect, rtol=1e-13)   # neighbour is element 1 (O)
    np.testing.assert_allclose(G[0, 0:32], 0.0)
    np.testing.assert_allclose(G[1, 0:32], expect, rtol=1e-13)


def test_angular_right_angle_closed_form():
    """theta = pi/2: the t=2 term (theta_t = pi/2) equals 2^(1-z) 2^z = 2, the
    t=0 and t=4 terms equal 2^(1-z), and t=6 (theta - theta_t = -pi) vanishes;
    a sign error in theta - theta_t swaps the t=2 and t=6 values."""
    sp = O.SymParams([1, 8])
    x = torch.tensor([[0.0, 0, 0], [1.0, 0, 0], [0, 1.0, 0]], dtype=torch.float64)
    G = O.sym_angular(x, np.array([0]), np.array([1]), np.array([2]), np.array([0]), 3, sp).numpy()
    fc = 0.5 * (math.cos(3.14159265359 * 1.0 / 3.1) + 1.0)
    rad = np.array([math.exp(-4.0 * (1.0 - 3.1 * s / 8) ** 2) for s in range(8)])
    g = G[0, :64].reshape(8, 8)
    np.testing.assert_allclose(g[2], 2.0 * rad * fc * fc, rtol=1e-12)
    np.testing.assert_allclose(g[0], 2.0 ** -7 * rad * fc * fc, rtol=1e-12)
    assert np.all(np.abs(g[6]) < 1e-30)
    # theta_t = pi (t=4): (1 + cos(pi/2))^8 2^-7 = 2^-7
    np.testing.assert_allclose(g[4], 2.0 ** -7 * rad * fc * fc, rtol=1e-10)
    assert np.all(G[0, 64:] == 0)


# ---------------------------------------------------------------- networks
def test_mlp_matches_torch_linear():
    rng = np.random.default_rng(0)
    from tensormol_amd.model import glorot_layers
    layers = glorot_layers(rng, [16, 8, 8, 1])
    Gx = rng.normal(size=(5, 16))
    y = O.mlp(torch.tensor(Gx), O._t(layers), 1.0).numpy()
    h = torch.tensor(Gx)
    for (Wm, b) in layers[:-1]:
        h = torch.nn.functional.softplus(torch.nn.functional.linear(h, torch.tensor(Wm).T, torch.tensor(b)))
    Wm, b = layers[-1]
    ref = torch.nn.functional.linear(h, torch.tensor(Wm).T, torch.tensor(b))[:, 0].numpy()
    np.testing.assert_allclose(y, ref, rtol=1e-13)
    # sharp activation: log(1+exp(100 z))/100
    z = torch.tensor([-3.0, -0.01, 0.0, 0.02, 5.0], dtype=torch.float64)
    np.testing.assert_allclose(O.activation(z, 100.0).numpy(),
                               np.log1p(np.exp(100 * z.numpy())) / 100, rtol=1e-12)


# --------------------------------------------------------------- Coulomb
def test_dsf_kernel_continuity_and_cutoff():
    c = O.dsf_consts(0.2)
    a, Rsr, Rlr, shift, ea = c
    r = torch.tensor([Rsr - 1e-7, Rsr + 1e-7, Rlr - 1e-6, Rlr + 1e-6], dtype=torch.float64,
                     requires_grad=True)
    K = O.dsf_kernel(r, c)
    (dK,) = torch.autograd.grad(K.sum(), r)
    assert abs(K[0].item() - K[1].item()) < 1e-8          # continuous at R_sr
    assert abs(dK[0].item() - dK[1].item()) < 1e-6        # C1 at R_sr
    assert abs(K[2].item()) < 1e-12 and K[3].item() == 0.0  # shifted to zero at R_lr
    assert abs(dK[2].item()) < 1e-9                       # shifted force
    # closed form of Util.DSF in the long-range region
    from scipy.special import erfc
    rr = 12.0
    XX = a * Rlr
    ZZ = erfc(XX) / Rlr
    YY = 1.1283791671 * a * math.exp(-XX * XX) / Rlr
    ref = erfc(a * rr) / rr - ZZ + (rr - Rlr) * (ZZ / Rlr + YY)
    assert abs(O.dsf_kernel(torch.tensor([rr], dtype=torch.float64), c).item() - ref) < 1e-14


def test_vdw_one_pair_closed_form():
    c6, r0 = O.vdw_params([8, 1])
    rA = 3.0
    rb = rA * O.BOHRPERA
    t64 = lambda v: torch.tensor([v], dtype=torch.float64)
    V = O.vdw_kernel(t64(rb), t64(c6[0]), t64(c6[1]), t64(r0[0]), t64(r0[1])).item()
    rv = rA * O.BOHRPERA ** 2
    t = rv / (4.6 * O.BOHRPERA)
    S = 1.0 if t > 1 else -t * t * (2 * t - 3)
    ref = -S * math.sqrt(c6[0] * c6[1]) / rv ** 6 / (1 + 6 * (rv / (r0[0] + r0[1])) ** -12)
    assert abs(V - ref) < 1e-12 * abs(ref)
    # C6 unit conversion: 0.70 J nm^6/mol for O (PhysicalData.py:26)
    assert abs(c6[0] - 0.70 * (O.BOHRPERA * 10) ** 6 / O.JOULEPERHARTREE) < 1e-9


# ------------------------------------------------------------ full energy
@pytest.fixture(scope="module")
def small_water():
    from tensormol_amd.model import BPEEModel
    Z, x = W.water_cluster(8, box=6.0, seed=1)
    model = BPEEModel([1, 8], dsf_alpha=0.2, seed=3, hidden=[32, 32, 32])
    sp = O.SymParams(model.eles)
    return Z, x, model, sp


def _E(x, Z, model, sp, **kw):
    return O.evaluate(x, Z, np.array([0, len(Z)]), sp, model.energy_nets, model.charge_nets,
                      dsf_alpha=model.dsf_alpha, **kw)


def test_forces_match_central_differences(small_water):
    Z, x, model, sp = small_water
    r = _E(x, Z, model, sp)
    L = O.build_lists(x, Z, np.array([0, len(Z)]), sp)
    rng = np.random.default_rng(0)
    h = 1e-5
    for _ in range(8):
        a, d = int(rng.integers(len(Z))), int(rng.integers(3))
        xp, xm = x.copy(), x.copy()
        xp[a, d] += h
        xm[a, d] -= h
        ep = _E(xp, Z, model, sp, lists=L, need_forces=False)["Etotal"][0]
        em = _E(xm, Z, model, sp, lists=L, need_forces=False)["Etotal"][0]
        fd = (ep - em) / (2 * h)
        assert abs(fd - r["gradient"][a, d]) < 1e-6 * max(1.0, abs(fd)), (a, d, fd, r["gradient"][a, d])


def test_translation_rotation_invariance(small_water):
    Z, x, model, sp = small_water
    r0 = _E(x, Z, model, sp)
    R = W.random_rotation(np.random.default_rng(9))
    x2 = x @ R.T + np.array([3.0, -1.0, 7.0])
    r1 = _E(x2, Z, model, sp)
    assert abs(r1["Etotal"][0] - r0["Etotal"][0]) < 1e-9 * abs(r0["Etotal"][0])
    np.testing.assert_allclose(r1["gradient"], r0["gradient"] @ R.T, atol=1e-10)
    np.testing.assert_allclose(r0["gradient"].sum(0), 0.0, atol=1e-10)
    torque = np.cross(x, r0["gradient"]).sum(0)
    np.testing.assert_allclose(torque, 0.0, atol=1e-9)


def test_permutation_invariance_and_neutrality(small_water):
    Z, x, model, sp = small_water
    r0 = _E(x, Z, model, sp)
    p = np.random.default_rng(2).permutation(len(Z))
    r1 = _E(x[p], Z[p], model, sp)
    assert abs(r1["Etotal"][0] - r0["Etotal"][0]) < 1e-9 * abs(r0["Etotal"][0])
    np.testing.assert_allclose(r1["gradient"], r0["gradient"][p], atol=1e-10)
    assert abs(r0["charge"].sum()) < 1e-12


def test_nve_energy_conservation(small_water):
    """VelocityVerletStep (SimpleMD.py:14-38) conserves E over 40 steps at 0.5 fs."""
    Z, x, model, sp = small_water
    m = O.atomic_mass_kg_per_mol(Z)
    v = W.maxwell_boltzmann(Z, 300.0, 0, m)

    def eandf(xx):
        r = _E(xx, Z, model, sp)
        return r["Etotal"][0], -O.JOULEPERHARTREE * r["gradient"]
    e0, f0 = eandf(x)
    a = 1e-10 * f0 / m[:, None]
    xx, vv = x.copy(), v.copy()
    etot = []
    for _ in range(40):
        xx, vv, a, e = O.vv_step(xx, vv, a, m, 0.5, eandf)
        ke = O.kinetic_energy(vv, m) * len(m) / O.JOULEPERHARTREE
        etot.append(e + ke)
    etot = np.array(etot)
    assert np.ptp(etot) < 2e-4 * len(Z)


# ------------------------------------------------------------ vdW: properties the mathematics fixes
def _vdw(rA, zi=8, zj=1):
    c6, r0 = O.vdw_params([zi, zj])
    t64 = lambda v: torch.tensor(np.atleast_1d(v), dtype=torch.float64)
    return O.vdw_kernel(t64(np.asarray(rA) * O.BOHRPERA), t64(c6[0]), t64(c6[1]), t64(r0[0]),
                        t64(r0[1])).numpy(), c6, r0


def test_vdw_large_r_limit():
    """Far beyond the switch and the damping radius the pair term is the bare
    dispersion -sqrt(C6_i C6_j) / r^6 (r in the kernel's own distance unit)."""
    for rA in (20.0, 40.0):
        V, c6, r0 = _vdw(rA)
        r = rA * O.BOHRPERA ** 2
        assert abs(V[0] * r ** 6 / -math.sqrt(c6[0] * c6[1]) - 1.0) < 1e-12


def test_vdw_vanishes_at_contact_and_is_symmetric():
    # as r -> 0 the switch starts as 3 t^2 and the damping turns r^-6 into
    # r^6 / (6 R0^12): V ~ r^8, so doubling a small r multiplies V by 2^8
    V, _, _ = _vdw([1e-5, 2e-5])
    assert abs(V[1] / V[0] / 256.0 - 1.0) < 1e-4 and V[0] < 0.0
    rs = np.linspace(0.5, 12.0, 40)
    np.testing.assert_array_equal(_vdw(rs, 8, 1)[0], _vdw(rs, 1, 8)[0])


def test_vdw_switch_ends_at_width_over_bohrpera():
    """The switch S(t) = -t^2 (2t - 3) ends (S = 1, S' = 0) at t = 1.  The
    reference scales the bohr distance by BOHRPERA once more inside the term
    (RawSymFunc.py:1377), so in the physical distance the switch ends at
    Poly_Width / BOHRPERA = 2.434 A, not at 4.6 A; beyond it the term equals
    the unswitched one, before it its magnitude is smaller, and V and dV/dr
    are continuous there."""
    a = 4.6 / O.BOHRPERA
    for rA in (a * 1.0001, 3.0, 4.0, 4.6):
        V, c6, r0 = _vdw(rA)
        r = rA * O.BOHRPERA ** 2
        bare = -math.sqrt(c6[0] * c6[1]) / r ** 6 / (1.0 + 6.0 * (r / (r0[0] + r0[1])) ** -12.0)
        assert abs(V[0] / bare - 1.0) < 1e-13, rA
    for rA in (0.5 * a, 0.9 * a, 0.999 * a):
        V, c6, r0 = _vdw(rA)
        r = rA * O.BOHRPERA ** 2
        bare = -math.sqrt(c6[0] * c6[1]) / r ** 6 / (1.0 + 6.0 * (r / (r0[0] + r0[1])) ** -12.0)
        assert abs(V[0]) < abs(bare) * (1 - 1e-7), rA
    h = 1e-6
    (vm, v0, vp), _, _ = _vdw([a - h, a, a + h])
    assert abs(vp - vm) < 20 * h * abs(v0)                        # continuous (|dV/dr| ~ 6|V|/r)
    dl = (v0 - vm) / h
    dr = (vp - v0) / h
    assert abs(dl - dr) < 1e-3 * abs(dr)                          # C1: S'(1) = 0


# ------------------------------------------------------------ FIRE (Bitzek et al. 2006)
def _bowl(x0, k=2e5):
    """E = k/2 |x - x0|^2 in J/mol; force J/mol/A."""
    return lambda x: -k * (x - x0)


def test_fire_converges_on_a_bowl():
    rng = np.random.default_rng(0)
    x0 = rng.normal(size=(5, 3))
    m = np.full(5, 0.012)
    force = _bowl(x0)
    x = x0 + rng.normal(0, 0.3, size=x0.shape)
    v = np.zeros_like(x)
    st = [1.0, 0.1, 0]
    for _ in range(300):
        x, v, st = O.fire_step(x, v, force(x), m, st, dtmax=10.0)
    assert np.abs(x - x0).max() < 1e-6


def test_fire_reset_growth_and_cap():
    m = np.full(2, 0.012)
    f = np.array([[1e4, 0.0, 0.0], [0.0, -1e4, 0.0]])
    # moving against the force (P < 0): v reset, dt * fdec, alpha reset
    x1, v1, st = O.fire_step(np.zeros((2, 3)), -f * 1e-9, f, m, [1.0, 0.05, 3], dtmax=10.0)
    assert st == [0.5, 0.1, 0]
    np.testing.assert_allclose(v1, 0.5 * 1e-10 * f / m[:, None], rtol=1e-15)
    # moving with the force: alpha and dt untouched for nmin steps, then dt * finc, alpha * falpha
    v = f * 1e-9
    st = [1.0, 0.1, 0]
    dts = []
    for _ in range(7):
        _, v, st = O.fire_step(np.zeros((2, 3)), v, f, m, st, dtmax=1.15)
        dts.append(st[0])
    np.testing.assert_allclose(dts, [1.0] * 5 + [1.1, 1.15])
    # a huge force: each atom moves exactly maxstep
    x2, _, _ = O.fire_step(np.zeros((2, 3)), np.zeros((2, 3)), f * 1e12, m, [1.0, 0.1, 0],
                           dtmax=10.0, maxstep=0.1)
    np.testing.assert_allclose(np.linalg.norm(x2, axis=1), 0.1, rtol=1e-14)
