"""Geometry optimisers.

GeomOptimizer is the drop-in of TensorMol/Simulations/Opt.py:17-75 with the
Polak-Ribiere conjugate gradient and golden-section line search of
TensorMol/Math/QuasiNewtonTools.py:350-464 and RemoveInvariantForce
(:551-577); it calls a user EnAndForce(x, DoForce) closure.

DeviceFIRE is the resident batched optimiser of the serving path (FIRE,
Bitzek et al., Phys. Rev. Lett. 97, 170201 (2006); the reference has no FIRE):
every molecule of a packed batch carries its own time step, mixing factor and
convergence flag on the device.
"""
import numpy as np
import torch

from . import _lib
from .containers import Mol
from .params import PARAMS, JOULEPERHARTREE, atomic_masses

_P = _lib.ptr
GOLDENRATIO = (np.sqrt(5.0) + 1.0) / 2.0


def InertiaTensor(x_, m_):
    I = np.zeros((3, 3))
    for i in range(len(m_)):
        I[0, 0] += m_[i] * (x_[i, 1] ** 2 + x_[i, 2] ** 2)
        I[1, 1] += m_[i] * (x_[i, 0] ** 2 + x_[i, 2] ** 2)
        I[2, 2] += m_[i] * (x_[i, 1] ** 2 + x_[i, 0] ** 2)
        I[0, 1] -= m_[i] * x_[i, 0] * x_[i, 1]
        I[0, 2] -= m_[i] * x_[i, 0] * x_[i, 2]
        I[1, 2] -= m_[i] * x_[i, 1] * x_[i, 2]
    I[1, 0], I[2, 0], I[2, 1] = I[0, 1], I[0, 2], I[1, 2]
    return I


def RemoveInvariantForce(x_, f_, m_):
    """Remove the centre-of-mass force and the rigid-rotation part (QuasiNewtonTools.py:551-577)."""
    if not PARAMS["RemoveInvariant"]:
        return f_
    m_ = np.asarray(m_, dtype=np.float64)
    fnet = np.sum(f_, axis=0)
    fnew = f_ - np.einsum("m,f->mf", m_, fnet) / np.sum(m_)
    torque = np.sum(np.cross(x_, fnew), axis=0)
    U, s, V = np.linalg.svd(InertiaTensor(x_, m_))
    s = np.where(np.abs(s) > 1e-7, 1.0 / np.where(s == 0, 1, s), 0.0)
    dwdt = (U @ np.diag(s) @ V) @ torque
    fcorr = np.zeros(f_.shape)
    fcorr[:, 0] = m_ * (-dwdt[2] * x_[:, 1] + dwdt[1] * x_[:, 2])
    fcorr[:, 1] = m_ * (dwdt[2] * x_[:, 0] - dwdt[0] * x_[:, 2])
    fcorr[:, 2] = m_ * (-dwdt[1] * x_[:, 0] + dwdt[0] * x_[:, 1])
    return fnew - fcorr


class ConjGradient:
    """Polak-Ribiere CG with a golden-section line search (QuasiNewtonTools.py:350-464)."""

    def __init__(self, f_, x0_, thresh_=0.0001):
        self.EForce = f_
        self.Energy = lambda x: self.EForce(x, False)
        self.xold = x0_.copy()
        self.e, self.gold = self.EForce(x0_)
        self.s = self.gold.copy()
        self.thresh = thresh_
        self.alpha = PARAMS["GSSearchAlpha"]

    def BetaPR(self, g):
        betapr = np.sum(g * (g - self.gold)) / np.sum(self.gold * self.gold)
        self.gold = g.copy()
        return max(0, betapr)

    def __call__(self, x0):
        e, g = self.EForce(x0)
        beta = self.BetaPR(g)
        self.s = g + beta * self.s
        self.xold = self.LineSearch(x0, self.s, self.thresh)
        return self.xold, e, g

    def LineSearch(self, x0_, p_, thresh=0.0001):
        a = x0_
        b = x0_ + self.alpha * p_
        c = b - (b - a) / GOLDENRATIO
        d = a + (b - a) / GOLDENRATIO
        fa, fb, fc, fd = self.Energy(a), self.Energy(b), self.Energy(c), self.Energy(d)
        rmsdist = 10.0
        while rmsdist > thresh:
            if fa < fc and fa < fd and fa < fb:
                if self.alpha > 0.00001:
                    self.alpha /= 1.8001
                else:
                    return a
                b = x0_ + self.alpha * p_
                c = b - (b - a) / GOLDENRATIO
                d = a + (b - a) / GOLDENRATIO
                fa, fb, fc, fd = self.Energy(a), self.Energy(b), self.Energy(c), self.Energy(d)
            elif fb < fc and fb < fd and fb < fa:
                if self.alpha < 100.0:
                    self.alpha *= 1.8
                b = x0_ + self.alpha * p_
                return (b + a) / 2
            elif fc < fd:
                b = d
                c = b - (b - a) / GOLDENRATIO
                d = a + (b - a) / GOLDENRATIO
                fb = fd
                fc, fd = self.Energy(c), self.Energy(d)
            else:
                a = c
                c = b - (b - a) / GOLDENRATIO
                d = a + (b - a) / GOLDENRATIO
                fa = fc
                fc, fd = self.Energy(c), self.Energy(d)
            rmsdist = np.sum(np.linalg.norm(a - b, axis=1)) / a.shape[0]
        return (b + a) / 2


class GeomOptimizer:
    """GeomOptimizer(f_).Opt(m_) (Opt.py:17-75)."""

    def __init__(self, f_):
        self.thresh = PARAMS["OptThresh"]
        self.max_opt_step = PARAMS["OptMaxCycles"]
        self.EnergyAndForce = f_
        self.m = None

    def WrappedEForce(self, x_, DoForce=True):
        if DoForce:
            energy, frc = self.EnergyAndForce(x_, DoForce)
            # the reference passes atomic numbers as the weights (Opt.py:39)
            frc = RemoveInvariantForce(x_, frc, self.m.atoms)
            frc /= JOULEPERHARTREE
            return energy, frc
        return self.EnergyAndForce(x_, False)

    def Opt(self, m_, filename="OptLog", Debug=False):
        m = Mol(m_.atoms, m_.coords)
        self.m = m
        rmsdisp, rmsgrad, step = 10.0, 10.0, 0
        prev_m = Mol(m.atoms, m.coords)
        CG = ConjGradient(self.WrappedEForce, m.coords)
        while step < self.max_opt_step and rmsgrad > self.thresh and (rmsdisp > 0.000001 or step < 5):
            prev_m = Mol(m.atoms, m.coords)
            m.coords, energy, frc = CG(m.coords)
            rmsgrad = np.sum(np.linalg.norm(frc, axis=1)) / m.coords.shape[0]
            rmsdisp = np.sum(np.linalg.norm(m.coords - prev_m.coords, axis=1)) / m.coords.shape[0]
            prev_m.properties["Step"] = step
            prev_m.properties["Energy"] = energy
            step += 1
        return prev_m


class DeviceFIRE:
    """Batched FIRE on the device over an engine workspace.

    Standard parameters (Bitzek et al. 2006): N_min 5, f_inc 1.1, f_dec 0.5,
    alpha_start 0.1, f_alpha 0.99; dt0 = 1 fs and dt_max = 10 dt0 (about
    ASE's defaults of 0.1 and 1.0 in its time unit of 10.18 fs).  Moves are capped at
    `maxstep` A per atom per iteration.  A molecule is converged when its
    maximum atomic force falls below fmax (Ha/A); converged molecules stop
    moving (their state is frozen on the device).
    """

    def __init__(self, engine, ws, Z, x0, dt0=1.0, fmax=1e-3, maxstep=0.1):
        self.eng = engine
        self.ws = ws
        dev = engine.device
        self.m = torch.tensor(atomic_masses(Z), device=dev)
        self.pos = torch.tensor(np.asarray(x0, dtype=np.float64), device=dev)
        self.v = torch.zeros_like(self.pos)
        self.dots = torch.zeros(ws.N, 3, dtype=torch.float64, device=dev)
        self.dots_sum = torch.zeros(ws.B, 3, dtype=torch.float64, device=dev)
        st = np.zeros((ws.B, 4))
        st[:, 0] = dt0
        st[:, 1] = 0.1
        st[:, 3] = 1.0
        self.state = torch.tensor(st, device=dev)
        self.mix = torch.zeros(ws.B, 3, dtype=torch.float64, device=dev)
        self.dt0, self.fmax, self.maxstep = dt0, fmax, maxstep

    def step(self):
        e, ws, L = self.eng, self.ws, self.eng.lib
        st = e._stream()
        e.run(ws, self.pos, check=False)
        L.tm_fire_dots(ws.N, _P(ws.force), _P(self.v), _P(self.dots), st)
        L.tm_mol_sum_f64(ws.B, _P(ws.mol_start), 3, _P(self.dots), 3, _P(self.dots_sum), 0,
                         _P(ws.red), ws.N, st)
        L.tm_fire_update(ws.N, ws.B, _P(ws.mol), 0, _P(self.dots_sum), _P(self.state), _P(self.mix),
                         _P(ws.force), _P(self.m), _P(self.v), _P(self.pos), 10.0 * self.dt0, 5, 1.1,
                         0.5, 0.1, 0.99, self.maxstep, st)

    def converged(self):
        """Per-molecule convergence (max |F| < fmax Ha/A); freezes converged
        molecules.  Synchronises."""
        g = self.ws.grad.cpu().numpy()
        fm = np.linalg.norm(g, axis=1)
        mx = np.zeros(self.ws.B)
        np.maximum.at(mx, self.ws.mol_np, fm)
        conv = mx < self.fmax
        st = self.state.cpu().numpy()
        st[conv, 3] = 0.0
        self.state.copy_(torch.tensor(st))
        return conv
