"""Molecular dynamics drivers.

Drop-in host-side drivers with the reference's signatures and units (A, fs,
kg/mol, forces J/mol/A), restating TensorMol/Simulations/SimpleMD.py:
VelocityVerletStep (:14-38), KineticEnergy (:40-51), Thermostat (:53-88),
NoseThermostat (:90-129), VelocityVerlet (:322-425, including its a = 0 first
step).  They call a user EnAndForce closure, as in the reference.

DeviceMD is the resident path: positions, velocities, accelerations and the
thermostat state of every molecule of a packed batch live in HBM, each step is
a fixed sequence of HIP launches (FP64 integrator kernels around one engine
evaluation) and the whole step can be captured into a HIP graph.
"""
import time

import numpy as np
import torch

from . import _lib
from .params import PARAMS, IDEALGASR, JOULEPERHARTREE, KJPERHARTREE, atomic_masses

_P = _lib.ptr


# ------------------------------------------------------------ host drivers
def VelocityVerletStep(f_, a_, x_, v_, m_, dt_, fande_=None):
    x = x_ + v_ * dt_ + 0.5 * a_ * dt_ * dt_
    e, f_x_ = 0.0, None
    if fande_ is None:
        f_x_ = f_(x)
    else:
        e, f_x_ = fande_(x)
    a = 1e-10 * np.einsum("ax,a->ax", f_x_, 1.0 / m_)
    v = v_ + 0.5 * (a_ + a) * dt_
    return x, v, a, e


def KineticEnergy(v_, m_):
    """Kinetic energy per atom in J/mol."""
    return 0.5 * np.dot(np.einsum("ia,ia->i", v_, v_) * 1e10, m_) / len(m_)


class Thermostat:
    """Velocity rescaling to PARAMS["MDTemp"] (SimpleMD.py:53-88)."""

    def __init__(self, m_, v_):
        self.N = len(m_)
        self.m = m_.copy()
        self.T = PARAMS["MDTemp"]
        self.Teff = 0.001
        self.kT = IDEALGASR * 1e-10 * self.T
        self.tau = 30 * PARAMS["MDdt"]
        self.name = "Rescaling"
        self.Rescale(v_)

    def step(self, f_, a_, x_, v_, m_, dt_, fande_=None):
        x = x_ + v_ * dt_ + 0.5 * a_ * dt_ * dt_
        e, f_x_ = 0.0, None
        if fande_ is None:
            f_x_ = f_(x)
        else:
            e, f_x_ = fande_(x)
        a = 1e-10 * np.einsum("ax,a->ax", f_x_, 1.0 / m_)
        v = v_ + 0.5 * (a_ + a) * dt_
        self.Teff = (2.0 / 3.0) * KineticEnergy(v, self.m) / IDEALGASR
        v *= np.sqrt(self.T / self.Teff)
        return x, v, a, e, f_x_

    def Rescale(self, v_):
        """Each atom rescaled to exactly T (SimpleMD.py:82-88)."""
        for i in range(self.N):
            Teff = (2.0 / (3.0 * IDEALGASR)) * 1e10 * 0.5 * self.m[i] * np.dot(v_[i], v_[i])
            if Teff != 0.0:
                v_[i] *= np.sqrt(self.T / Teff)


class NoseThermostat(Thermostat):
    """Single-variable Nose-Hoover (SimpleMD.py:90-129)."""

    def __init__(self, m_, v_):
        self.m = m_.copy()
        self.N = len(m_)
        self.T = PARAMS["MDTemp"]
        self.eta = 0.0
        self.name = "Nose"
        self.Rescale(v_)

    def step(self, f_, a_, x_, v_, m_, dt_, fande_=None, frc_=True):
        self.kT = IDEALGASR * 1e-10 * self.T
        self.tau = 20.0 * PARAMS["MDdt"] * self.N
        self.Q = self.kT * self.tau * self.tau
        x = x_ + v_ * dt_ + 0.5 * (a_ - self.eta * v_) * dt_ * dt_
        vdto2 = v_ + 0.5 * (a_ - self.eta * v_) * dt_
        e, f_x_ = 0.0, None
        if fande_ is None:
            f_x_ = f_(x)
        else:
            e, f_x_ = fande_(x)
        a = 1e-10 * np.einsum("ax,a->ax", f_x_, 1.0 / m_)
        ke = 0.5 * np.dot(np.einsum("ia,ia->i", v_, v_), m_)
        etadto2 = self.eta + (dt_ / (2.0 * self.Q)) * (ke - ((3.0 * self.N + 1) / 2.0) * self.kT)
        kedto2 = 0.5 * np.dot(np.einsum("ia,ia->i", vdto2, vdto2), m_)
        self.eta = etadto2 + (dt_ / (2.0 * self.Q)) * (kedto2 - ((3.0 * self.N + 1) / 2.0) * self.kT)
        v = (vdto2 + (dt_ / 2.0) * a) / (1 + (dt_ / 2.0) * self.eta)
        if frc_:
            return x, v, a, e, f_x_
        return x, v, a, e


class VelocityVerlet:
    """VelocityVerlet(f_, g0_, name_, EandF_, cellsize_).Prop() (SimpleMD.py:322-425)."""

    def __init__(self, f_, g0_, name_="", EandF_=None, cellsize_=None):
        self.name = name_
        self.cellsize = cellsize_
        self.maxstep = PARAMS["MDMaxStep"]
        self.T = PARAMS["MDTemp"]
        self.dt = PARAMS["MDdt"]
        self.ForceFunction = f_
        self.EnergyAndForce = EandF_
        self.EPot0 = 0.0
        if EandF_ is not None:
            self.EPot0, self.f0 = self.EnergyAndForce(g0_.coords)
        self.EPot = self.EPot0
        self.t = 0.0
        self.KE = 0.0
        self.atoms = g0_.atoms.copy()
        self.m = atomic_masses(self.atoms)
        self.natoms = len(self.atoms)
        self.x = g0_.coords.copy()
        self.v = np.zeros(self.x.shape)
        self.a = np.zeros(self.x.shape)   # the reference starts from a = 0 (SimpleMD.py:359)
        self.md_log = None
        if PARAMS["MDV0"] == "Random":
            self.v = np.random.randn(*self.x.shape)
            Thermostat(self.m, self.v)
        self.Tstat = None
        if PARAMS["MDThermostat"] == "Rescaling":
            self.Tstat = Thermostat(self.m, self.v)
        elif PARAMS["MDThermostat"] == "Nose":
            self.Tstat = NoseThermostat(self.m, self.v)

    def Prop(self):
        step = 0
        self.md_log = np.zeros((self.maxstep, 7))
        while step < self.maxstep:
            self.t = step * self.dt
            if PARAMS["MDThermostat"] is None:
                self.x, self.v, self.a, self.EPot = VelocityVerletStep(
                    self.ForceFunction, self.a, self.x, self.v, self.m, self.dt, self.EnergyAndForce)
            else:
                self.x, self.v, self.a, self.EPot, self.force = self.Tstat.step(
                    self.ForceFunction, self.a, self.x, self.v, self.m, self.dt, self.EnergyAndForce)
            if self.cellsize is not None:
                self.x = np.mod(self.x, self.cellsize)
            self.md_log[step, 0] = self.t
            self.md_log[step, 4] = self.KE
            self.md_log[step, 5] = self.EPot
            self.md_log[step, 6] = self.KE + (self.EPot - self.EPot0) * JOULEPERHARTREE
            self.KE = KineticEnergy(self.v, self.m)
            step += 1
        return


# ------------------------------------------------------------ device path
class DeviceMD:
    """Resident batched integrator over an engine workspace.

    kind: "nve" (Velocity Verlet) or "nvt" (Nose-Hoover with one thermostat
    variable per molecule, tau = 20 dt N as in SimpleMD.py:109-111).  The first
    step uses the forces of the initial positions (a = F/m).
    """

    def __init__(self, engine, ws, Z, x0, v0=None, dt=0.5, kind="nve", T=300.0, seed=0):
        self.eng = engine
        self.ws = ws
        self.dt = float(dt)
        self.kind = kind
        dev = engine.device
        m = atomic_masses(Z)
        self.m = torch.tensor(m, device=dev)
        if v0 is None:
            from .workloads import maxwell_boltzmann
            v0 = maxwell_boltzmann(Z, T, seed, m)
        self.pos = torch.tensor(np.asarray(x0, dtype=np.float64), device=dev)
        self.v = torch.tensor(np.asarray(v0, dtype=np.float64), device=dev)
        self.a = torch.zeros_like(self.v)
        self.vh = torch.zeros_like(self.v)
        self.ke = torch.zeros(ws.N, dtype=torch.float64, device=dev)
        self.ke2 = torch.zeros(ws.N, 2, dtype=torch.float64, device=dev)
        self.ke2sum = torch.zeros(ws.B, 2, dtype=torch.float64, device=dev)
        self.eta = torch.zeros(ws.B, dtype=torch.float64, device=dev)
        n = np.diff(ws.ms_np).astype(np.float64)
        kT = IDEALGASR * 1e-10 * T
        tau = 20.0 * self.dt * n
        self.kT = torch.tensor(np.full(ws.B, kT), device=dev)
        self.Q = torch.tensor(kT * tau * tau, device=dev)
        self.graph = None

    def start(self):
        """Forces at the initial positions -> accelerations."""
        e, ws = self.eng, self.ws
        e.run(ws, self.pos)
        e.lib.tm_vv_kick(ws.N, 0.0, _P(ws.force), _P(self.m), _P(self.v), _P(self.a), _P(self.ke),
                         e._stream())

    def step(self):
        e, ws, L = self.eng, self.ws, self.eng.lib
        st = e._stream()
        if self.kind == "nve":
            L.tm_vv_drift(ws.N, self.dt, _P(self.v), _P(self.a), _P(self.pos), st)
        else:
            L.tm_nh_drift(ws.N, self.dt, _P(ws.mol), _P(self.eta), _P(self.m), _P(self.pos),
                          _P(self.v), _P(self.a), _P(self.vh), _P(self.ke2), st)
        e.run(ws, self.pos, check=False)
        if self.kind == "nve":
            L.tm_vv_kick(ws.N, self.dt, _P(ws.force), _P(self.m), _P(self.v), _P(self.a), _P(self.ke),
                         st)
        else:
            L.tm_mol_sum_f64(ws.B, _P(ws.mol_start), 2, _P(self.ke2), 2, _P(self.ke2sum), 0,
                             _P(ws.red), ws.N, st)
            L.tm_nh_eta(ws.B, _P(ws.mol_start), self.dt, _P(self.kT), _P(self.Q), _P(self.ke2sum),
                        _P(self.eta), st)
            L.tm_nh_kick(ws.N, self.dt, _P(ws.mol), _P(self.eta), _P(ws.force), _P(self.m),
                         _P(self.vh), _P(self.v), _P(self.a), _P(self.ke), st)

    def run_steps(self, n):
        """n steps, leaving the state of n calls to step() bit for bit.  For
        NVE the kick of every step but the last and the drift of the step after
        it run inside the force gather's launch (tm_force_gather_vv), so a step
        costs two launches fewer; the status flags are cleared once and
        accumulate over the n steps.  Capturable into one HIP graph."""
        if self.kind != "nve" or n <= 0:
            for _ in range(n):
                self.step()
            return
        e, ws, L = self.eng, self.ws, self.eng.lib
        st = e._stream()
        L.tm_vv_drift(ws.N, self.dt, _P(self.v), _P(self.a), _P(self.pos), st)
        for k in range(n):
            e.build_lists(ws, self.pos, clear_flags=k == 0)
            e.descriptors(ws, self.pos)
            e.networks(ws)
            e.electrostatics(ws, self.pos)
            if k + 1 < n:
                e.adjoint(ws, self.pos, vv=(self.dt, self.m, self.v, self.a, self.ke, self.pos))
                e.reductions(ws)
            else:
                e.adjoint(ws, self.pos)
                e.reductions(ws)
                L.tm_vv_kick(ws.N, self.dt, _P(ws.force), _P(self.m), _P(self.v), _P(self.a),
                             _P(self.ke), st)

    def capture(self, stream):
        """Capture one step into a HIP graph on `stream` (replay with .replay())."""
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=stream):
            self.step()
        return self.graph

    def energies(self):
        """(potential Ha, kinetic Ha) per molecule, synchronising."""
        pe = self.eng.energy_tensor(self.ws).cpu().numpy()
        ke = np.zeros(self.ws.B)
        k = (self.ke.cpu().numpy() * 1e10 / JOULEPERHARTREE)
        np.add.at(ke, self.ws.mol_np, k)
        return pe, ke
