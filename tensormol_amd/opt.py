"""Geometry optimisers.

GeomOptimizer(f).Opt(m) has the call shape of TensorMol/Simulations/Opt.py:17-75,
with the conjugate-gradient minimiser of TensorMol/Math/QuasiNewtonTools.py:350-464
and the invariant-force projection of :551-577.  It is written from this
specification of the reference's behaviour:

- Forces: the user closure f(x, True) -> (E Ha, F J/mol/A) and f(x, False) -> E.
  The optimiser projects the net force and the rigid-rotation torque out of F
  (weights: the atomic numbers, which is what the reference passes, Opt.py:39)
  when PARAMS["RemoveInvariant"] is set, and divides by JOULEPERHARTREE.
- Conjugate gradient (Polak-Ribiere with restart at beta < 0): g is the force
  (downhill).  At x: beta = max(0, g.(g - g_prev) / g_prev.g_prev), the
  direction s = g + beta s, and the next point is the line-search minimum
  along s from x.  The first direction is the force at the starting point.
- Line search: golden-section search of E on the segment [x, x + alpha s] with
  interior points at 1/phi from each end, until the mean per-atom distance
  between the segment ends is at most 1e-4 A (at least one pass).  If E is
  lowest at the start, alpha shrinks by 1.8001 and the bracket restarts (below
  alpha = 1e-5 the start is returned: the reference keeps re-evaluating the
  same bracket there); if lowest at the far end, alpha grows by 1.8 (up to 100)
  and the midpoint of the enlarged segment is returned without further search.
  alpha starts at PARAMS["GSSearchAlpha"] and persists between iterations.
- Opt(m): iterate while step < OptMaxCycles, the mean per-atom |F| exceeds
  OptThresh, and the mean per-atom displacement of the last step exceeds 1e-6
  A (or fewer than 5 steps were taken).  Each iteration appends the geometry
  it started from, with its Step and Energy, to results/<filename>.xyz.  Opt
  returns that geometry of the last iteration.

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
INV_GOLDEN = 2.0 / (1.0 + np.sqrt(5.0))   # 1 / phi


def InertiaTensor(x_, m_):
    """sum_i m_i (|r_i|^2 1 - r_i r_i^T) about the origin."""
    x = np.asarray(x_, dtype=np.float64)
    m = np.asarray(m_, dtype=np.float64)
    r2 = np.einsum("i,ia,ia->", m, x, x)
    return r2 * np.eye(3) - np.einsum("i,ia,ib->ab", m, x, x)


def RemoveInvariantForce(x_, f_, m_):
    """Project the centre-of-mass force and the rigid-rotation part out of the
    forces f_ (weights m_), QuasiNewtonTools.py:551-577: subtract m_i F_net /
    sum m, then m_i (w x r_i) with w = I^+ tau, where tau is the torque of the
    remaining forces about the origin and I^+ the pseudo-inverse of the
    inertia tensor (singular values below 1e-7 dropped)."""
    if not PARAMS["RemoveInvariant"]:
        return f_
    m = np.asarray(m_, dtype=np.float64)
    f = f_ - np.outer(m, f_.sum(axis=0)) / m.sum()
    tau = np.cross(x_, f).sum(axis=0)
    w = _pinv(InertiaTensor(x_, m)) @ tau
    return f - m[:, None] * np.cross(w, x_)


def _pinv(M, cut=1e-7):
    """Pseudo-inverse from the SVD, singular values <= cut treated as zero."""
    U, sv, Vt = np.linalg.svd(M)
    inv = np.zeros_like(sv)
    keep = np.abs(sv) > cut
    inv[keep] = 1.0 / sv[keep]
    return (Vt.T * inv) @ U.T


def _mean_atom_norm(d):
    """Mean over atoms of the Euclidean norm of each row."""
    return float(np.linalg.norm(d, axis=1).sum() / d.shape[0])


def _golden_points(lo, hi):
    """The two golden-section interior points of the segment [lo, hi]."""
    step = (hi - lo) * INV_GOLDEN
    return hi - step, lo + step


class ConjGradient:
    """Polak-Ribiere conjugate gradient driven by a force closure
    f(x, DoForce=True) -> (E, force) / E."""

    def __init__(self, f_, x0_, thresh_=0.0001):
        self.EForce = f_
        self.Energy = lambda x: self.EForce(x, False)
        self.x0 = x0_.copy()
        self.xold = x0_.copy()
        self.e, self.gold = self.EForce(x0_)
        self.s = self.gold.copy()
        self.thresh = thresh_
        self.alpha = PARAMS["GSSearchAlpha"]

    def BetaPR(self, g):
        """Polak-Ribiere beta, clipped at 0; remembers g for the next call."""
        prev = self.gold
        beta = float(np.vdot(g, g - prev) / np.vdot(prev, prev))
        self.gold = g.copy()
        return max(0.0, beta)

    def __call__(self, x0):
        """One CG iteration from x0: returns (next point, E(x0), force(x0))."""
        e, g = self.EForce(x0)
        self.s = g + self.BetaPR(g) * self.s
        self.xold = self.LineSearch(x0, self.s, self.thresh)
        return self.xold, e, g

    def LineSearch(self, x0_, p_, thresh=0.0001):
        """Golden-section minimum of the energy along p_ from x0_."""
        def fresh_bracket():
            lo, hi = x0_, x0_ + self.alpha * p_
            c, d = _golden_points(lo, hi)
            return lo, hi, c, d, self.Energy(lo), self.Energy(hi), self.Energy(c), self.Energy(d)

        a, b, c, d, fa, fb, fc, fd = fresh_bracket()
        while True:
            if fa < fb and fa < fc and fa < fd:          # minimum at the start: step too long
                if self.alpha <= 0.00001:
                    return a
                self.alpha /= 1.8001
                a, b, c, d, fa, fb, fc, fd = fresh_bracket()
            elif fb < fa and fb < fc and fb < fd:        # minimum at the far end: step too short
                if self.alpha < 100.0:
                    self.alpha *= 1.8
                return x0_ + 0.5 * self.alpha * p_
            else:                                        # shrink towards the lower interior point
                if fc < fd:
                    b, fb = d, fd
                else:
                    a, fa = c, fc
                c, d = _golden_points(a, b)
                fc, fd = self.Energy(c), self.Energy(d)
            if _mean_atom_norm(a - b) <= thresh:
                return 0.5 * (a + b)


class GeomOptimizer:
    """GeomOptimizer(f_).Opt(m_): conjugate-gradient relaxation of one molecule
    with a user closure f_(x, DoForce)."""

    def __init__(self, f_):
        self.thresh = PARAMS["OptThresh"]
        self.maxstep = PARAMS["OptMaxStep"]
        self.fscale = PARAMS["OptStepSize"]
        self.momentum = PARAMS["OptMomentum"]
        self.momentum_decay = PARAMS["OptMomentumDecay"]
        self.max_opt_step = PARAMS["OptMaxCycles"]
        self.step = self.maxstep
        self.EnergyAndForce = f_
        self.m = None

    def WrappedEForce(self, x_, DoForce=True):
        """(E Ha, invariant-projected force Ha/A), or E alone."""
        if not DoForce:
            return self.EnergyAndForce(x_, False)
        energy, force = self.EnergyAndForce(x_, True)
        force = RemoveInvariantForce(x_, force, self.m.atoms) / JOULEPERHARTREE
        return energy, force

    def Opt(self, m_, filename="OptLog", Debug=False):
        m = Mol(m_.atoms, m_.coords)
        self.m = m
        cg = ConjGradient(self.WrappedEForce, m.coords)
        last = Mol(m.atoms, m.coords)
        step, rmsgrad, rmsdisp = 0, np.inf, np.inf
        while step < self.max_opt_step and rmsgrad > self.thresh and (rmsdisp > 0.000001 or step < 5):
            last = Mol(m.atoms, m.coords)
            m.coords, energy, force = cg(m.coords)
            rmsgrad = _mean_atom_norm(force)
            rmsdisp = _mean_atom_norm(m.coords - last.coords)
            last.properties["Step"] = step
            last.properties["Energy"] = energy
            last.WriteXYZfile(PARAMS["results_dir"], filename, "a", True)
            step += 1
        return last


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
