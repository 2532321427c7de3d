"""Periodic drivers: the lattice, the periodic force wrapper and periodic MD.

Call shapes of TensorMol/ForceModifiers/Periodic.py (Lattice :12-180,
PeriodicForce :277-400) and TensorMol/Simulations/PeriodicMD.py
(PeriodicVelocityVerletStep :21-42, PeriodicNoseThermostat :44-75,
PeriodicVelocityVerlet :77-146), written from this specification:

- Lattice(latvec): rows are the lattice vectors a, b, c.  The centre is
  (a + b + c) / 2; latticeMinDiameter is twice the smallest distance from the
  centre to the 14 reference points (the origin, the 7 other corners, and
  the 6 points (a+b)/2, (b+c)/2, (a+c)/2, (a+b)/2 + c, (b+c)/2 + a,
  (a+c)/2 + b).  InLat gives fractional coordinates x L^T (L L^T)^-1, FromLat
  the inverse; ModuloLattice wraps fractional coordinates into [0, 1).
  TessLattice(atoms, coords, rng) returns the real atoms followed by their
  images shifted by i a + j b + k c for (i, j, k) in [-n, n]^3 in
  lexicographic order (i slowest), the origin skipped, with n = 1, or
  int(rng / latticeMinDiameter) + 1 when rng exceeds that diameter.
- PeriodicForce(mol, lat): the molecule is centred in the cell and wrapped;
  BindForce(f, rng) registers a local force f(z, x, nreal[, DoForce]) with its
  range; calling the wrapper with primitive coordinates wraps them, tessellates
  over the largest bound range, sums the energies of every local force and
  returns (E, F on the nreal real atoms).
- Periodic Velocity Verlet and Nose-Hoover steps: the SimpleMD steps with the
  drifted positions wrapped into the cell before the force call.
  PeriodicVelocityVerlet.Prop logs the kinetic energy of the current step
  (taken before the step) and writes a trajectory frame every step when
  PARAMS["MDLogTrajectory"] is set.

The BP+EE evaluator behind a PeriodicForce is the manager's
EvalBPDirectEEUpdateSinglePeriodic, which reads the lattice back from the
tessellated images and evaluates only the real atoms, under the minimum image
(every edge must be at least twice the largest cutoff).  Its forces are the
exact gradient of the periodic energy; the reference's tessellated graph drops
the derivatives through image positions (TFMolManage.py:1353, Periodic.py:396).
"""
import os

import numpy as np

from .containers import Mol
from .md import (NoseThermostat, VelocityVerlet, KineticEnergy, _accel, _mv2, _energy_force)
from .params import PARAMS, AVOCONST, IDEALGASR, JOULEPERHARTREE, atomic_masses


class Lattice:
    def __init__(self, latvec_):
        self.lattice = np.array(latvec_, dtype=np.float64).reshape(3, 3)
        a, b, c = self.lattice
        self.latticeCenter = 0.5 * (a + b + c)
        d = np.linalg.norm(self.LatticeFacePoints() - self.latticeCenter, axis=1)
        self.latticeMinDiameter = 2.0 * float(d.min())
        self.lp = np.array([np.zeros(3), a, b, c, a + b, a + c, b + c, a + b + c])
        self.ntess = 1

    def LatticeFacePoints(self):
        """Corners, the origin and six face / edge midpoints (14 x 3)."""
        a, b, c = self.lattice
        return np.array([a, b, c, a + b, a + c, b + c, a + b + c, np.zeros(3),
                         0.5 * (a + b), 0.5 * (c + b), 0.5 * (a + c), 0.5 * (a + b) + c,
                         0.5 * (c + b) + a, 0.5 * (a + c) + b])

    def InLat(self, crds):
        """Cartesian -> lattice (fractional) coordinates."""
        L = self.lattice
        return np.asarray(crds) @ L.T @ np.linalg.inv(L @ L.T)

    def FromLat(self, crds):
        return np.asarray(crds) @ self.lattice

    def ModuloLattice(self, crds):
        """Coordinates wrapped into the primitive cell (fractional part in [0, 1))."""
        f = np.fmod(self.InLat(crds), 1.0)
        f[f < 0.0] += 1.0
        return self.FromLat(f)

    def CenteredInLattice(self, mol):
        m = Mol(mol.atoms, self.ModuloLattice(mol.coords - mol.Center() + self.latticeCenter))
        m.properties["Lattice"] = self.lattice.copy()
        return m

    def _shells(self, rng_):
        if rng_ > self.latticeMinDiameter:
            self.ntess = int(rng_ / self.latticeMinDiameter) + 1
        else:
            self.ntess = 1
        return self.ntess

    def TessLattice(self, atoms_, coords_, rng_):
        """Real atoms, then (2n+1)^3 - 1 shifted copies (see the module notes)."""
        n = self._shells(rng_)
        shifts = image_shifts(n)
        natom = len(atoms_)
        z = np.tile(np.asarray(atoms_, dtype=np.uint8), len(shifts) + 1)
        x = np.empty(((len(shifts) + 1) * natom, 3))
        x[:natom] = coords_
        x[natom:] = (np.asarray(coords_)[None] + (shifts @ self.lattice)[:, None, :]).reshape(-1, 3)
        return z, x


def image_shifts(n):
    """(i, j, k) in [-n, n]^3, lexicographic with i slowest, origin excluded."""
    r = np.arange(-n, n + 1)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), axis=-1).reshape(-1, 3)
    return g[np.any(g != 0, axis=1)].astype(np.float64)


def lattice_from_images(coords, nreal, tol=1e-8):
    """Lattice vectors (3 x 3) of a coordinate array laid out as
    Lattice.TessLattice lays it out (nreal real atoms, then the images), or
    None when the array is not such a tessellation."""
    x = np.asarray(coords, dtype=np.float64)
    if nreal <= 0 or len(x) % nreal:
        return None
    nimg = len(x) // nreal
    n = int(round((nimg ** (1.0 / 3.0) - 1) / 2))
    if n < 1 or (2 * n + 1) ** 3 != nimg:
        return None
    shifts = image_shifts(n)
    d = x[nreal:].reshape(nimg - 1, nreal, 3) - x[None, :nreal]
    disp = d[:, 0, :]
    scale = max(1.0, float(np.abs(x).max()))
    if np.abs(d - disp[:, None, :]).max() > tol * scale:
        return None                       # the copies are not rigid shifts of the real atoms
    L, *_ = np.linalg.lstsq(shifts, disp, rcond=None)
    if np.abs(shifts @ L - disp).max() > tol * scale:
        return None
    return L


def orthorhombic_edges(lat, tol=1e-9):
    """Box edges of a diagonal lattice; raises for a skewed cell."""
    L = np.asarray(lat, dtype=np.float64)
    if L.ndim == 1:
        return L.reshape(3)
    off = L - np.diag(np.diag(L))
    if np.abs(off).max() > tol * np.abs(L).max():
        raise NotImplementedError("only orthorhombic cells are supported by the minimum-image "
                                  "kernels (the lattice has off-diagonal components)")
    return np.diag(L).copy()


class LocalForce:
    """A bound local force f(z, x, nreal[, DoForce]) with its range (A)."""

    def __init__(self, f_, rng_=5.0, NeedsTriples_=False):
        self.range = rng_
        self.func = f_
        self.NeedsTriples = NeedsTriples_

    def __call__(self, z, x, NZ, DoForce=True):
        return self.func(z, x, NZ, DoForce)


class PeriodicForce:
    def __init__(self, pm_, lat_):
        self.lattice = Lattice(lat_)
        self.NL = None
        self.mol0 = self.lattice.CenteredInLattice(pm_)
        self.atoms = self.mol0.atoms.copy()
        self.natoms = self.mol0.NAtoms()
        self.natomsReal = pm_.NAtoms()
        self.maxrng = 0.0
        self.LocalForces = []
        self.lastx = np.zeros(pm_.coords.shape)
        self.nlthresh = 0.05

    def ReLattice(self, lat_):
        self.lattice = Lattice(lat_)

    def Density(self):
        """Density of the bulk in g/cm^3."""
        grams = atomic_masses(self.mol0.atoms).sum() * 1000.0 / AVOCONST
        return grams / (np.linalg.det(self.lattice.lattice) * 1e-24)

    def AdjustLattice(self, x_, lat0_, latp_):
        """Coordinates x_ in lattice lat0_ carried to the same fractional
        positions in latp_."""
        return Lattice(lat0_).InLat(x_) @ np.asarray(latp_)

    def BindForce(self, lf_, rng_):
        self.LocalForces.append(LocalForce(lf_, rng_))

    def _tess(self, x_, lattice=None):
        lat = self.lattice if lattice is None else lattice
        return lat.TessLattice(self.atoms, lat.ModuloLattice(x_), self.maxrng)

    def __call__(self, x_, DoForce=True):
        """(energy per cell, forces on the real atoms) for primitive coordinates x_."""
        if self.maxrng == 0.0:
            self.maxrng = max(f.range for f in self.LocalForces)
        z, x = self._tess(x_)
        energy = 0.0
        force = np.zeros((self.natomsReal, 3))
        for f in self.LocalForces:
            if DoForce:
                e, frc = f(z, x, self.natomsReal)
                force += np.asarray(frc)[:self.natomsReal]
            else:
                e = f(z, x, self.natomsReal, DoForce)
            energy += np.sum(e)
        return energy, force

    def LatticeStep(self, x_):
        """Greedy lattice relaxation: try +/- PARAMS["OptLatticeStep"] on each
        of the nine lattice components with the last bound force and keep any
        change that lowers the energy by more than 1e-5; repeat until none
        does.  Halves the step when nothing moved.  Returns the wrapped
        coordinates."""
        xx = np.array(x_, dtype=np.float64)
        e, _ = self(xx)
        dlat = PARAMS["OptLatticeStep"]
        moved_any, moved = False, True
        while moved:
            moved = False
            for i in range(3):
                for j in range(3):
                    for sgn in (1.0, -1.0):
                        trial = self.lattice.lattice.copy()
                        trial[i, j] += sgn * dlat
                        lat = Lattice(trial)
                        xt = lat.ModuloLattice(xx)
                        z, x = lat.TessLattice(self.atoms, xt, self.maxrng)
                        et, _ = self.LocalForces[-1](z, x, self.natomsReal)
                        if et < e and abs(e - et) > 0.00001:
                            e, xx = et, xt
                            self.ReLattice(trial)
                            moved = moved_any = True
        if not moved_any and PARAMS["OptLatticeStep"] > 0.001:
            PARAMS["OptLatticeStep"] = PARAMS["OptLatticeStep"] / 2.0
        return xx

    def Save(self, x_, name_="PMol"):
        m = Mol(self.atoms, x_)
        m.properties["Lattice"] = np.array_str(self.lattice.lattice.flatten())
        m.WriteXYZfile(PARAMS["results_dir"], name_, "w", True)


def PeriodicVelocityVerletStep(pf_, a_, x_, v_, m_, dt_):
    """Velocity Verlet step with the drifted positions wrapped into the cell."""
    x = pf_.lattice.ModuloLattice(x_ + v_ * dt_ + 0.5 * a_ * dt_ ** 2)
    energy, force = pf_(x)
    a = _accel(force, m_)
    return x, v_ + 0.5 * dt_ * (a_ + a), a, energy


class PeriodicNoseThermostat(NoseThermostat):
    def step(self, pf_, a_, x_, v_, m_, dt_):
        """Nose-Hoover step of NoseThermostat with the drifted positions wrapped."""
        self._couple()
        g_kT = 0.5 * (3.0 * self.N + 1.0) * self.kT
        rate = dt_ / (2.0 * self.Q)
        a_damped = a_ - self.eta * v_
        x = pf_.lattice.ModuloLattice(x_ + v_ * dt_ + 0.5 * a_damped * dt_ ** 2)
        v_half = v_ + 0.5 * a_damped * dt_
        energy, force = pf_(x)
        a = _accel(force, m_)
        eta_half = self.eta + rate * (0.5 * _mv2(v_, m_) - g_kT)
        self.eta = eta_half + rate * (0.5 * _mv2(v_half, m_) - g_kT)
        v = (v_half + 0.5 * dt_ * a) / (1.0 + 0.5 * dt_ * self.eta)
        return x, v, a, energy


class PeriodicVelocityVerlet(VelocityVerlet):
    """PeriodicVelocityVerlet(PF, name, v0).Prop(): MD of the PeriodicForce's
    molecule; PARAMS["MDThermostat"] == "Nose" selects the periodic thermostat."""

    def __init__(self, Force_, name_="PdicMD", v0_=None):
        self.PForce = Force_
        thermo = PARAMS["MDThermostat"]
        try:
            PARAMS["MDThermostat"] = None     # the thermostat is chosen below
            VelocityVerlet.__init__(self, None, Force_.mol0, name_, Force_.__call__)
        finally:
            PARAMS["MDThermostat"] = thermo
        if v0_ is not None:
            self.v = np.array(v0_, dtype=np.float64)
        self.Tstat = PeriodicNoseThermostat(self.m, self.v) if thermo == "Nose" else None

    def Density(self):
        return self.PForce.Density()

    def WriteTrajectory(self):
        frame = Mol(self.atoms, self.x)
        frame.properties.update(Lattice=self.PForce.lattice.lattice.copy(), Time=self.t,
                                KineticEnergy=self.KE, PotEnergy=self.EPot)
        frame.WriteXYZfile(PARAMS["results_dir"], "MDTrajectory" + self.name, "a", True)

    def Prop(self):
        self.md_log = np.zeros((self.maxstep, 7))
        for step in range(self.maxstep):
            self.t = step * self.dt
            self.KE = KineticEnergy(self.v, self.m)
            if self.Tstat is None:
                self.x, self.v, self.a, self.EPot = PeriodicVelocityVerletStep(
                    self.PForce, self.a, self.x, self.v, self.m, self.dt)
            else:
                self.x, self.v, self.a, self.EPot = self.Tstat.step(
                    self.PForce, self.a, self.x, self.v, self.m, self.dt)
            row = self.md_log[step]
            row[0], row[4], row[5] = self.t, self.KE, self.EPot
            row[6] = self.KE + (self.EPot - self.EPot0) * JOULEPERHARTREE
            if PARAMS["MDLogTrajectory"]:
                self.WriteTrajectory()
            if step % 500 == 0:
                self._save_log()
        if self.maxstep:
            self._save_log()
