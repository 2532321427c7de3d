"""Molecular dynamics drivers.

Host drivers with the call shapes of TensorMol/Simulations/SimpleMD.py
(VelocityVerletStep :14-38, KineticEnergy :40-51, Thermostat :53-88,
NoseThermostat :90-129, VelocityVerlet :322-425).  They call the user's
EnAndForce closure exactly as the reference does, and are written from the
following specification of its behaviour.

Units: positions A, time fs, masses kg/mol (PhysicalData ATOMICMASSES),
forces J/mol/A.  Acceleration a = 1e-10 F / m in A/fs^2 (1 m/s^2 = 1e-10
A/fs^2); kinetic energies in kg/mol A^2/fs^2 = 1e10 J/mol.

- Velocity Verlet step: drift x' = x + v dt + a dt^2 / 2, force at x',
  a' = 1e-10 F'/m, v' = v + (a + a') dt / 2; returns (x', v', a', E(x')), E = 0
  when only a force closure is given.
- KineticEnergy: the mean over atoms of m |v|^2 / 2, in J/mol.
- Rescale (the velocity initialiser): every atom's velocity is scaled on its
  own so that its temperature m |v|^2 1e10 / (3 R) is exactly T; atoms at rest
  are left alone.
- Rescaling thermostat step: a Velocity Verlet step, then every velocity is
  scaled by sqrt(T / T_eff), T_eff = (2/3) KineticEnergy / R.
- Nose-Hoover step (one thermostat variable eta): kT = R 1e-10 T,
  tau = 20 dt_MD N, Q = kT tau^2, target g kT with g = (3N + 1) / 2,
  K(u) = sum m |u|^2 / 2:
      x'  = x + v dt + (a - eta v) dt^2 / 2
      vh  = v + (a - eta v) dt / 2
      eta <- eta + dt/(2Q) (K(v) - g kT) + dt/(2Q) (K(vh) - g kT)
      v'  = (vh + a' dt / 2) / (1 + eta dt / 2)
- VelocityVerlet(f, g0, name, EandF, cellsize).Prop(): the initial potential
  EPot0 comes from EandF(g0.coords) when given.  The state starts from
  a = 0, so the first step ignores the initial forces.  Velocities start at 0,
  or standard-normal then Rescale'd for PARAMS["MDV0"] == "Random".  The
  thermostat comes from PARAMS["MDThermostat"] (None, "Rescaling", "Nose").
  Each of PARAMS["MDMaxStep"] steps wraps positions into [0, cellsize) when
  a cell size is given.  It appends the row [t, 0, 0, 0, KE, EPot,
  KE + (EPot - EPot0) JOULEPERHARTREE] to md_log, where KE is the per-atom
  kinetic energy of the previous step (0 on the first), then updates KE and
  the running mean / variance of EPot.  It writes the trajectory every third
  step when PARAMS["MDLogTrajectory"] is set and saves md_log every 500 steps.
  (The reference's Rescaling thermostat returns four values where Prop unpacks
  five; here it returns five so that the thermostat is usable.)

DeviceMD is the resident path: positions, velocities, accelerations and the
thermostat state of every molecule of a packed batch live in HBM, each step is
a fixed sequence of HIP launches (FP64 integrator kernels around one engine
evaluation) and the whole step can be captured into a HIP graph.
"""
import os
import time

import numpy as np
import torch

from . import _lib
from .containers import Mol
from .params import PARAMS, IDEALGASR, JOULEPERHARTREE, KJPERHARTREE, atomic_masses
from .statistics import OnlineEstimator

_P = _lib.ptr


# ------------------------------------------------------------ host drivers
def _energy_force(x, force_fn, energy_force_fn):
    """(E, F) at x from whichever closure the caller supplied."""
    if energy_force_fn is None:
        return 0.0, force_fn(x)
    return energy_force_fn(x)


def _accel(force, masses):
    """J/mol/A over kg/mol -> A/fs^2."""
    return 1e-10 * np.asarray(force) / masses[:, None]


def _mv2(v, masses):
    """sum_i m_i |v_i|^2 in kg/mol A^2/fs^2 (twice the kinetic energy)."""
    return float(masses @ np.square(v).sum(axis=1))


def VelocityVerletStep(f_, a_, x_, v_, m_, dt_, fande_=None):
    """One Velocity Verlet step; returns (x, v, a, E) at the new positions."""
    x = x_ + v_ * dt_ + 0.5 * a_ * dt_ ** 2
    energy, force = _energy_force(x, f_, fande_)
    a = _accel(force, m_)
    return x, v_ + 0.5 * dt_ * (a_ + a), a, energy


def KineticEnergy(v_, m_):
    """Mean kinetic energy per atom, J/mol."""
    return 0.5e10 * _mv2(v_, m_) / len(m_)


def _atom_temperatures(v, masses):
    """Per-atom temperature m |v|^2 1e10 / (3 R), in K."""
    return masses * np.square(v).sum(axis=1) * 1e10 / (3.0 * IDEALGASR)


class Thermostat:
    """Velocity-rescaling thermostat at PARAMS["MDTemp"]; constructing it
    rescales v_ in place atom by atom."""

    name = "Rescaling"

    def __init__(self, m_, v_):
        self.m = np.array(m_, dtype=np.float64)
        self.N = len(self.m)
        self.T = PARAMS["MDTemp"]
        self.kT = IDEALGASR * 1e-10 * self.T
        self.tau = 30 * PARAMS["MDdt"]
        self.Teff = 0.001
        self.Rescale(v_)

    def Rescale(self, v_):
        """Scale each atom's velocity (in place) to temperature T."""
        Ti = _atom_temperatures(v_, self.m)
        moving = Ti != 0.0
        v_[moving] *= np.sqrt(self.T / Ti[moving])[:, None]

    def step(self, f_, a_, x_, v_, m_, dt_, fande_=None):
        x, v, a, energy = VelocityVerletStep(f_, a_, x_, v_, m_, dt_, fande_)
        self.Teff = (2.0 / 3.0) * KineticEnergy(v, self.m) / IDEALGASR
        v = v * np.sqrt(self.T / self.Teff)
        return x, v, a, energy, None


class NoseThermostat(Thermostat):
    """Nose-Hoover thermostat with one variable eta (tau = 20 dt N)."""

    name = "Nose"

    def __init__(self, m_, v_):
        self.m = np.array(m_, dtype=np.float64)
        self.N = len(self.m)
        self.T = PARAMS["MDTemp"]
        self.eta = 0.0
        self.Rescale(v_)

    def _couple(self):
        """kT, tau and Q from the current T, N and PARAMS["MDdt"]."""
        self.kT = IDEALGASR * 1e-10 * self.T
        self.tau = 20.0 * PARAMS["MDdt"] * self.N
        self.Q = self.kT * self.tau ** 2

    def step(self, f_, a_, x_, v_, m_, dt_, fande_=None, frc_=True):
        self._couple()
        g_kT = 0.5 * (3.0 * self.N + 1.0) * self.kT
        rate = dt_ / (2.0 * self.Q)
        a_damped = a_ - self.eta * v_
        x = x_ + v_ * dt_ + 0.5 * a_damped * dt_ ** 2
        v_half = v_ + 0.5 * a_damped * dt_
        energy, force = _energy_force(x, f_, fande_)
        a = _accel(force, m_)
        eta_half = self.eta + rate * (0.5 * _mv2(v_, m_) - g_kT)
        self.eta = eta_half + rate * (0.5 * _mv2(v_half, m_) - g_kT)
        v = (v_half + 0.5 * dt_ * a) / (1.0 + 0.5 * dt_ * self.eta)
        return (x, v, a, energy, force) if frc_ else (x, v, a, energy)


_THERMOSTATS = {"Rescaling": Thermostat, "Nose": NoseThermostat}


class VelocityVerlet:
    """VelocityVerlet(f_, g0_, name_, EandF_, cellsize_): .Prop() runs
    PARAMS["MDMaxStep"] steps from the molecule g0_ and fills md_log."""

    def __init__(self, f_, g0_, name_="", EandF_=None, cellsize_=None):
        self.name = name_
        self.cellsize = cellsize_
        self.maxstep = int(PARAMS["MDMaxStep"])
        self.T = PARAMS["MDTemp"]
        self.dt = PARAMS["MDdt"]
        self.ForceFunction = f_
        self.EnergyAndForce = EandF_
        self.atoms = g0_.atoms.copy()
        self.natoms = len(self.atoms)
        self.m = atomic_masses(self.atoms)
        self.x = np.array(g0_.coords, dtype=np.float64)
        self.EPot0 = 0.0
        if EandF_ is not None:
            self.EPot0, self.f0 = EandF_(self.x)
        self.EPot = self.EPot0
        self.EnergyStat = OnlineEstimator(self.EPot0)
        self.t = 0.0
        self.KE = 0.0
        self.a = np.zeros_like(self.x)
        self.v = np.zeros_like(self.x)
        if PARAMS["MDV0"] == "Random":
            self.v = np.random.standard_normal(self.x.shape)
            Thermostat(self.m, self.v)
        kind = PARAMS["MDThermostat"]
        if kind is not None and kind not in _THERMOSTATS:
            raise ValueError(f"MDThermostat {kind!r}: use None, 'Rescaling' or 'Nose'")
        self.Tstat = None if kind is None else _THERMOSTATS[kind](self.m, self.v)
        self.md_log = None

    def _advance(self):
        """One integrator step of the configured kind; updates x, v, a, EPot."""
        if self.Tstat is None:
            self.x, self.v, self.a, self.EPot = VelocityVerletStep(
                self.ForceFunction, self.a, self.x, self.v, self.m, self.dt, self.EnergyAndForce)
        else:
            self.x, self.v, self.a, self.EPot, self.force = self.Tstat.step(
                self.ForceFunction, self.a, self.x, self.v, self.m, self.dt, self.EnergyAndForce)
        if self.cellsize is not None:
            self.x = np.mod(self.x, self.cellsize)

    def WriteTrajectory(self):
        """Append the current frame to results/MDTrajectory<name>.xyz."""
        frame = Mol(self.atoms, self.x)
        frame.properties.update(Time=self.t, KineticEnergy=self.KE, PotEnergy=self.EPot)
        frame.WriteXYZfile(PARAMS["results_dir"], "MDTrajectory" + self.name)

    def _save_log(self):
        os.makedirs(PARAMS["results_dir"], exist_ok=True)
        np.savetxt(os.path.join(PARAMS["results_dir"], "MDLog" + self.name + ".txt"), self.md_log)

    def Prop(self):
        self.md_log = np.zeros((self.maxstep, 7))
        for step in range(self.maxstep):
            self.t = step * self.dt
            self._advance()
            row = self.md_log[step]
            row[0] = self.t
            row[4] = self.KE
            row[5] = self.EPot
            row[6] = self.KE + (self.EPot - self.EPot0) * JOULEPERHARTREE
            self.EnergyStat(self.EPot)
            self.KE = KineticEnergy(self.v, self.m)
            if PARAMS["MDLogTrajectory"] and step % 3 == 0:
                self.WriteTrajectory()
            if step % 500 == 0:
                self._save_log()
        if self.maxstep:
            self._save_log()


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

    def step(self, check=True):
        """One step.  With check the status flags are read afterwards (a
        synchronisation) and an overflowed neighbour row, Coulomb range buffer
        or a non-finite coordinate raises, as BPEEEngine.check_flags does;
        check=False leaves the step asynchronous (and capturable)."""
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
        if check:
            e.check_flags(ws)

    def run_steps(self, n, check=True):
        """n steps, leaving the state of n calls to step() bit for bit.  For
        NVE the kick of every step but the last and the drift of the step after
        it run inside the force gather's launch (tm_force_gather_vv), so a step
        costs two launches fewer; the status flags are cleared once and
        accumulate over the n steps, and with check they are read once at the
        end (raising as step() does).  With check=False it is capturable into
        one HIP graph; call check_flags() after the replays."""
        if self.kind != "nve" or n <= 0:
            for k in range(n):
                self.step(check=check and k == n - 1)
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
        if check:
            e.check_flags(ws)

    def check_flags(self):
        """Raise if any step since the flags were last cleared overflowed or
        saw a non-finite coordinate (synchronises)."""
        return self.eng.check_flags(self.ws)

    def capture(self, stream):
        """Capture one step into a HIP graph on `stream` (replay with .replay();
        call check_flags() after replays)."""
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=stream):
            self.step(check=False)
        return self.graph

    def energies(self):
        """(potential Ha, kinetic Ha) per molecule, synchronising."""
        pe = self.eng.energy_tensor(self.ws).cpu().numpy()
        ke = np.zeros(self.ws.B)
        k = (self.ke.cpu().numpy() * 1e10 / JOULEPERHARTREE)
        np.add.at(ke, self.ws.mol_np, k)
        return pe, ke
