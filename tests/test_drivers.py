"""Host MD / optimiser drivers (tensormol_amd.md, tensormol_amd.opt) and the
oracle's integrators, on the CPU.

The drivers take a user EnAndForce closure, so they are exercised here with a
smooth, translation- and rotation-invariant spring potential whose energy and
forces are exact.  Pins:
  - the oracle's Nose-Hoover step reduces to its Velocity Verlet step when the
    thermostat mass Q goes to infinity, and conserves the extended energy
    K + U + Q eta^2 + (3N+1) kT int(eta dt) up to an error that vanishes with
    dt (first order: the eta update uses K at t and t + dt/2, so halving dt
    halves the drift); a wrong target (3N instead of 3N+1), a dropped eta half
    update or a missing friction denominator leave a drift that does not
    shrink with dt (checked by planting each in a copy of the step);
  - the host drivers follow the oracle step for step (VelocityVerletStep,
    NoseThermostat.step, VelocityVerlet.Prop with its a = 0 first step);
  - Thermostat.Rescale gives every atom exactly the target temperature;
  - OnlineEstimator equals numpy's mean and sample variance;
  - RemoveInvariantForce leaves no net force and no torque; GeomOptimizer
    relaxes the springs to their rest lengths.
"""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from tensormol_amd import md as MD
from tensormol_amd import opt as OPT
from tensormol_amd.containers import Mol
from tensormol_amd.params import PARAMS, IDEALGASR, JOULEPERHARTREE, atomic_masses
from tensormol_amd.statistics import OnlineEstimator

# ---------------------------------------------------------------- potential
ATOMS = np.array([6, 1, 1, 8, 7, 1], dtype=np.uint8)
X0 = np.array([[0.0, 0.0, 0.0], [1.09, 0.0, 0.0], [-0.4, 1.0, 0.1], [-0.5, -0.8, 0.9],
               [0.3, -1.3, -0.9], [1.2, -1.6, -1.2]])
BONDS = [(0, 1), (0, 2), (0, 3), (0, 4), (4, 5), (1, 2), (2, 3), (3, 4), (1, 4), (2, 5), (1, 5),
         (3, 5)]
K_SPRING = 0.25   # Ha / A^2


def _rest(x):
    return {b: float(np.linalg.norm(x[b[0]] - x[b[1]])) for b in BONDS}


REST = _rest(X0)


def spring_energy_grad(x):
    """E (Ha) and dE/dx (Ha/A) of harmonic springs at their X0 lengths."""
    e = 0.0
    g = np.zeros_like(x)
    for (i, j), r0 in REST.items():
        d = x[i] - x[j]
        r = math.sqrt(d @ d)
        e += 0.5 * K_SPRING * (r - r0) ** 2
        gij = K_SPRING * (r - r0) * d / r
        g[i] += gij
        g[j] -= gij
    return e, g


def enandforce(x):
    """The reference's EnAndForce contract: (E Ha, F J/mol/A)."""
    e, g = spring_energy_grad(x)
    return e, -JOULEPERHARTREE * g


def distorted(seed=0, s=0.08):
    return X0 + np.random.default_rng(seed).normal(0.0, s, X0.shape)


def test_spring_forces_are_the_gradient():
    x = distorted(1)
    _, g = spring_energy_grad(x)
    h = 1e-6
    for a in range(len(x)):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, d] += h
            xm[a, d] -= h
            fd = (spring_energy_grad(xp)[0] - spring_energy_grad(xm)[0]) / (2 * h)
            assert abs(fd - g[a, d]) < 1e-7


# ------------------------------------------------------------ oracle pins
def _mb(m, T, seed):
    rng = np.random.default_rng(seed)
    # A/fs: sigma^2 = R T / m * 1e-10 (kg/mol A^2/fs^2 per J/mol)
    return rng.normal(size=(len(m), 3)) * np.sqrt(IDEALGASR * T * 1e-10 / m)[:, None]


def test_nose_step_large_q_reduces_to_vv():
    m = atomic_masses(ATOMS)
    x, v = distorted(2), _mb(m, 300.0, 2)
    a = 1e-10 * enandforce(x)[1] / m[:, None]
    xa, va, aa, eta = x.copy(), v.copy(), a.copy(), 0.0
    xb, vb, ab = x.copy(), v.copy(), a.copy()
    for _ in range(25):
        # mddt 1e30 makes tau, and Q = kT tau^2, astronomically large
        xa, va, aa, _, eta = O.nose_step(xa, va, aa, m, 0.5, enandforce, eta, 300.0, 1e30)
        xb, vb, ab, _ = O.vv_step(xb, vb, ab, m, 0.5, enandforce)
    assert abs(eta) < 1e-40
    np.testing.assert_allclose(xa, xb, rtol=0, atol=1e-13)
    np.testing.assert_allclose(va, vb, rtol=0, atol=1e-14)


def _nh_extended_drift(dt, nsteps, T=300.0):
    """Spread of K + U + Q' eta^2 / 2 + N_f kT int eta dt (Q' = 2Q, N_f = 3N+1,
    the constant of motion of d eta/dt = (K - (3N+1) kT / 2) / Q), in
    kg/mol A^2/fs^2, along oracle Nose-Hoover steps with tau = 20 dt_MD N."""
    m = atomic_masses(ATOMS)
    N = len(m)
    mddt = 0.05   # a short coupling time, so eta moves a lot within the run
    kT = IDEALGASR * 1e-10 * T
    Q = kT * (20.0 * mddt * N) ** 2
    x, v = distorted(3, 0.1), _mb(m, 2000.0, 3)
    a = 1e-10 * enandforce(x)[1] / m[:, None]
    eta, integ = 0.0, 0.0
    u = lambda xx: spring_energy_grad(xx)[0] * JOULEPERHARTREE * 1e-10
    ke = lambda vv: 0.5 * float(m @ (vv * vv).sum(1))
    H = []
    etas = []
    for _ in range(nsteps):
        eta_old = eta
        x, v, a, _, eta = O.nose_step(x, v, a, m, dt, enandforce, eta, T, mddt)
        integ += 0.5 * (eta_old + eta) * dt
        H.append(ke(v) + u(x) + Q * eta * eta + (3 * N + 1) * kT * integ)
        etas.append(eta)
    return np.ptp(H), np.ptp(etas), kT


def test_nose_step_conserves_extended_energy():
    d1, eta_range, kT = _nh_extended_drift(0.05, 800)
    d2, _, _ = _nh_extended_drift(0.025, 1600)
    assert eta_range > 0.1             # the thermostat variable is really moving
    assert d2 < 1.5 * kT, (d2, kT)
    assert d1 / d2 > 1.7, (d1, d2)     # the error vanishes with dt (first order)


# ------------------------------------------------------------ host drivers
def test_velocity_verlet_step_matches_oracle():
    m = atomic_masses(ATOMS)
    x, v = distorted(4), _mb(m, 300.0, 4)
    a = 1e-10 * enandforce(x)[1] / m[:, None]
    r_host = MD.VelocityVerletStep(None, a, x, v, m, 0.5, enandforce)
    r_orc = O.vv_step(x, v, a, m, 0.5, enandforce)
    for p, q in zip(r_host, r_orc):
        np.testing.assert_allclose(p, q, rtol=1e-14, atol=1e-18)
    # force-only closure: E reported as 0
    r2 = MD.VelocityVerletStep(lambda y: enandforce(y)[1], a, x, v, m, 0.5)
    assert r2[3] == 0.0
    np.testing.assert_allclose(r2[0], r_orc[0], rtol=1e-14)


def test_kinetic_energy_per_atom():
    m = atomic_masses(ATOMS)
    v = _mb(m, 300.0, 5)
    ref = sum(0.5 * m[i] * (v[i] @ v[i]) * 1e10 for i in range(len(m))) / len(m)
    assert abs(MD.KineticEnergy(v, m) - ref) < 1e-12 * ref


def test_rescale_gives_each_atom_the_target_temperature(monkeypatch):
    monkeypatch.setitem(PARAMS, "MDTemp", 450.0)
    m = atomic_masses(ATOMS)
    v = np.random.default_rng(6).standard_normal((len(m), 3))
    v[2] = 0.0                                # an atom at rest stays at rest
    MD.Thermostat(m, v)
    Ti = m * (v * v).sum(1) * 1e10 / (3 * IDEALGASR)
    np.testing.assert_allclose(np.delete(Ti, 2), 450.0, rtol=1e-12)
    assert np.all(v[2] == 0.0)


def test_nose_thermostat_matches_oracle(monkeypatch):
    monkeypatch.setitem(PARAMS, "MDTemp", 300.0)
    monkeypatch.setitem(PARAMS, "MDdt", 0.05)
    m = atomic_masses(ATOMS)
    x, v = distorted(7, 0.1), _mb(m, 900.0, 7)
    nh = MD.NoseThermostat(m, v.copy())       # rescales its own copy only
    a = 1e-10 * enandforce(x)[1] / m[:, None]
    xh, vh, ah = x.copy(), v.copy(), a.copy()
    xo, vo, ao, eta = x.copy(), v.copy(), a.copy(), 0.0
    for _ in range(30):
        xh, vh, ah, eh, fh = nh.step(None, ah, xh, vh, m, 0.5, enandforce)
        xo, vo, ao, eo, eta = O.nose_step(xo, vo, ao, m, 0.5, enandforce, eta, 300.0, 0.05)
        assert abs(eh - eo) < 1e-12
    assert abs(nh.eta - eta) < 1e-12 * max(1.0, abs(eta))
    np.testing.assert_allclose(xh, xo, rtol=0, atol=1e-11)
    np.testing.assert_allclose(vh, vo, rtol=0, atol=1e-13)
    assert nh.eta != 0.0
    # frc_=False: four values
    assert len(nh.step(None, ah, xh, vh, m, 0.5, enandforce, frc_=False)) == 4


def test_velocity_verlet_prop(monkeypatch, tmp_path):
    """Prop(): a = 0 on the first step, md_log columns, trajectory every third
    step, the log file, and the state equal to the oracle's steps."""
    monkeypatch.setitem(PARAMS, "MDMaxStep", 10)
    monkeypatch.setitem(PARAMS, "MDdt", 0.5)
    monkeypatch.setitem(PARAMS, "MDV0", None)
    monkeypatch.setitem(PARAMS, "MDThermostat", None)
    monkeypatch.setitem(PARAMS, "MDLogTrajectory", True)
    monkeypatch.setitem(PARAMS, "results_dir", str(tmp_path) + os.sep)
    x0 = distorted(8, 0.1)
    vv = MD.VelocityVerlet(None, Mol(ATOMS, x0), "t", enandforce)
    assert np.all(vv.v == 0.0) and np.all(vv.a == 0.0)
    vv.Prop()
    m = atomic_masses(ATOMS)
    x, v, a = x0.copy(), np.zeros_like(x0), np.zeros_like(x0)
    kes, eps = [], []
    for _ in range(10):
        x, v, a, e = O.vv_step(x, v, a, m, 0.5, enandforce)
        kes.append(O.kinetic_energy(v, m))
        eps.append(e)
    np.testing.assert_allclose(vv.x, x, rtol=0, atol=1e-12)
    np.testing.assert_allclose(vv.v, v, rtol=0, atol=1e-14)
    log = vv.md_log
    np.testing.assert_allclose(log[:, 0], 0.5 * np.arange(10))
    np.testing.assert_allclose(log[:, 5], eps, rtol=1e-12)
    # column 4 holds the previous step's kinetic energy (0 before the first step)
    np.testing.assert_allclose(log[:, 4], [0.0] + kes[:-1], rtol=1e-12)
    e0 = enandforce(x0)[0]
    np.testing.assert_allclose(log[:, 6], log[:, 4] + (log[:, 5] - e0) * JOULEPERHARTREE, rtol=1e-12)
    traj = open(tmp_path / "MDTrajectoryt.xyz").read()
    assert traj.count("Comment:") == 4                 # steps 0, 3, 6, 9
    np.testing.assert_allclose(np.loadtxt(tmp_path / "MDLogt.txt"), log, rtol=1e-12)
    mean, var = vv.EnergyStat.mean, vv.EnergyStat.m2 / (vv.EnergyStat.n - 1)
    allE = np.array([e0] + eps)
    assert abs(mean - allE.mean()) < 1e-12 and abs(var - allE.var(ddof=1)) < 1e-12


def test_velocity_verlet_prop_nose_and_wrap(monkeypatch, tmp_path):
    monkeypatch.setitem(PARAMS, "MDMaxStep", 6)
    monkeypatch.setitem(PARAMS, "MDdt", 0.5)
    monkeypatch.setitem(PARAMS, "MDV0", "Random")
    monkeypatch.setitem(PARAMS, "MDThermostat", "Nose")
    monkeypatch.setitem(PARAMS, "MDLogTrajectory", False)
    monkeypatch.setitem(PARAMS, "results_dir", str(tmp_path) + os.sep)
    np.random.seed(11)
    vv = MD.VelocityVerlet(None, Mol(ATOMS, distorted(9) + 5.0), "n", enandforce, cellsize_=6.0)
    m = atomic_masses(ATOMS)
    Ti = m * (vv.v ** 2).sum(1) * 1e10 / (3 * IDEALGASR)
    np.testing.assert_allclose(Ti, PARAMS["MDTemp"], rtol=1e-12)   # "Random" init, per-atom rescale
    vv.Prop()
    assert np.all((vv.x >= 0.0) & (vv.x < 6.0))
    assert isinstance(vv.Tstat, MD.NoseThermostat) and vv.Tstat.eta != 0.0
    with pytest.raises(ValueError):
        monkeypatch.setitem(PARAMS, "MDThermostat", "Andersen")
        MD.VelocityVerlet(None, Mol(ATOMS, X0), "x", enandforce)


def test_online_estimator():
    xs = np.random.default_rng(12).normal(3.0, 2.0, 50)
    est = OnlineEstimator(xs[0])
    for k in range(1, 50):
        mean, var = est(xs[k])
        assert abs(mean - xs[:k + 1].mean()) < 1e-12
        assert abs(var - xs[:k + 1].var(ddof=1)) < 1e-10


# ------------------------------------------------------------ optimiser
def test_remove_invariant_force(monkeypatch):
    monkeypatch.setitem(PARAMS, "RemoveInvariant", True)
    w = ATOMS.astype(np.float64)
    x = distorted(13)
    x = x - (w[:, None] * x).sum(0) / w.sum()       # weighted centre at the origin
    f = np.random.default_rng(13).normal(size=x.shape)
    g = OPT.RemoveInvariantForce(x, f, w)
    np.testing.assert_allclose(g.sum(0), 0.0, atol=1e-12)
    np.testing.assert_allclose(np.cross(x, g).sum(0), 0.0, atol=1e-12)
    # an already invariant force (internal springs) is unchanged
    fi = -spring_energy_grad(x)[1]
    np.testing.assert_allclose(OPT.RemoveInvariantForce(x, fi, w), fi, atol=1e-12)
    monkeypatch.setitem(PARAMS, "RemoveInvariant", False)
    assert OPT.RemoveInvariantForce(x, f, w) is f


def test_inertia_tensor_closed_form():
    x = np.array([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0]])
    I = OPT.InertiaTensor(x, np.array([2.0, 3.0]))
    np.testing.assert_allclose(I, np.diag([12.0, 2.0, 14.0]))


def test_geom_optimizer_relaxes_springs(monkeypatch, tmp_path):
    monkeypatch.setitem(PARAMS, "results_dir", str(tmp_path) + os.sep)
    monkeypatch.setitem(PARAMS, "OptMaxCycles", 200)
    monkeypatch.setitem(PARAMS, "OptThresh", 1e-5)

    def f(x, DoForce=True):
        e, fr = enandforce(x)
        return (e, fr) if DoForce else e
    m0 = Mol(ATOMS, distorted(14, 0.12))
    e_start = f(m0.coords, False)
    out = GeomOptimizerRun = OPT.GeomOptimizer(f).Opt(m0, filename="o")
    assert e_start > 1e-3
    assert out.properties["Energy"] < 1e-6 * e_start
    rest = _rest(out.coords)
    for b in BONDS:
        assert abs(rest[b] - REST[b]) < 2e-3
    log = open(tmp_path / "o.xyz").read()
    assert log.count("Comment:") == out.properties["Step"] + 1
    energies = [float(l.split(";;;Energy ")[1].split(";;;")[0]) for l in log.splitlines()
                if ";;;Energy " in l]
    assert all(b <= a + 1e-12 for a, b in zip(energies, energies[1:]))   # monotone descent
