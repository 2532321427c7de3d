"""Continuous-batching E+F server for a stream of molecule requests.

Request kinds: single-point energy + forces ("sp"), Nose-Hoover NVT segments
("nvt", per-molecule thermostat, SimpleMD.py:90-129 semantics) and FIRE
optimisations ("opt").  The reference has no batching: its only request loop
is the one-at-a-time i-PI client (TMIPIinterface.py:19-61), and a new molecule
size rebuilds its TensorFlow graph (TFMolInstanceDirect.py:5694-5699).

Every iteration of Server.run:
  1. admits requests whose arrival time has passed, while the packed batch
     stays within the atom capacity derived from an HBM budget;
  2. re-packs the active requests into kind-contiguous groups [nvt | opt | sp]
     (state of continuing requests moves with on-device gathers; new requests'
     coordinates are uploaded once);
  3. runs one E+F launch sequence over all atoms, then the NVT integrator
     kernels on the nvt range and the FIRE kernels on the opt range;
  4. retires finished requests (single points after one evaluation, NVT after
     their step count, FIRE once the molecule's RMS atomic force is below
     fire_fmax Ha/A, or at the step limit), copying
     their results out and recording the latency.
Per-request state never leaves the device while the request is active, and
every per-molecule quantity (descriptors, networks, Coulomb, reductions,
integrators) depends only on the molecule's own atoms, so a request gives the
same result alone and inside any mix.
"""
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .params import IDEALGASR, JOULEPERHARTREE, atomic_masses
from .workloads import STREAM_KINDS, maxwell_boltzmann, request_stream

_P = _lib.ptr
KIND_ORDER = {"nvt": 0, "opt": 1, "sp": 2}


@dataclass
class Request:
    rid: int
    Z: np.ndarray
    xyz: np.ndarray
    kind: str               # "sp" | "nvt" | "opt"
    nsteps: int             # NVT steps or FIRE step limit (1 for sp)
    t_arrival: float        # seconds after the stream start
    seed: int = 0
    # results
    energy: float = None
    forces: np.ndarray = None
    positions: np.ndarray = None
    steps_done: int = 0
    t_admit: float = None
    t_done: float = None
    converged: bool = None
    error: str = None       # set when the request failed (its molecule only)
    # server bookkeeping: atom offset and molecule slot in the current packed batch
    _off: int = field(default=-1, repr=False)
    _mol: int = field(default=-1, repr=False)


def make_stream(n=4096, rate=2000.0, lo=8, hi=256, seed=3, kinds=STREAM_KINDS, workers=None):
    """The seeded request stream of workloads.request_stream as Request objects."""
    return [Request(r, Z, x, kind, steps, t, seed=sd)
            for r, (Z, x, kind, steps, t, sd) in enumerate(
                request_stream(n, rate, lo, hi, seed, kinds, workers))]


class Server:
    """Continuous batching over one engine workspace of fixed capacity."""

    def __init__(self, engine, max_atoms=65536, max_mols=8192, hbm_budget_bytes=None, T=300.0,
                 dt=0.5, fire_dt0=1.0, fire_fmax=5e-3, fire_maxstep=0.1):
        self.eng = engine
        per_atom = self.bytes_per_atom()
        if hbm_budget_bytes is not None:
            max_atoms = min(max_atoms, int(hbm_budget_bytes // per_atom))
        self.max_atoms = int(max_atoms)
        self.max_mols = int(max_mols)
        self.ws = engine.alloc(self.max_atoms, self.max_mols)
        self.T, self.dt = T, dt
        self.fire = (fire_dt0, fire_fmax, fire_maxstep)
        dev = engine.device
        f64 = dict(dtype=torch.float64, device=dev)
        n, b = self.max_atoms, self.max_mols
        self.pos = torch.zeros(n, 3, **f64)
        self.v = torch.zeros(n, 3, **f64)
        self.a = torch.zeros(n, 3, **f64)
        self.vh = torch.zeros(n, 3, **f64)
        self.m = torch.zeros(n, **f64)
        self.ke = torch.zeros(n, **f64)
        self.ke2 = torch.zeros(n, 2, **f64)
        self.dots = torch.zeros(n, 3, **f64)
        self.eta = torch.zeros(b, **f64)
        self.kT = torch.zeros(b, **f64)
        self.Q = torch.ones(b, **f64)
        self.ke2sum = torch.zeros(b, 2, **f64)
        self.fstate = torch.zeros(b, 4, **f64)
        self.mix = torch.zeros(b, 3, **f64)
        self.dsum = torch.zeros(b, 3, **f64)
        self.active = []

    def bytes_per_atom(self):
        """Device bytes of one atom's share of the workspace and server state."""
        e = self.eng
        ws = 4 * e.D * 3 + e.ldr * (4 + 4 + 16) + e.lda * 4 + 4 * 2 * (e.nele + 1) + 200
        return ws + 8 * (3 * 4 + 2 + 2 + 3)

    # ---------------------------------------------------------------- packing
    def _repack(self, new):
        """Order active + new requests as [nvt | opt | sp], move device state."""
        old = self.active
        reqs = sorted(old + new, key=lambda r: (KIND_ORDER[r.kind], r.rid))
        sizes = np.array([len(r.Z) for r in reqs], dtype=np.int64)
        ms = np.concatenate([[0], np.cumsum(sizes)])
        N, B = int(ms[-1]), len(reqs)
        dev = self.eng.device
        # gather index for continuing requests (old offset -> new offset)
        src, dst = [], []
        for k, r in enumerate(reqs):
            if r._off >= 0:
                src.append(np.arange(r._off, r._off + len(r.Z)))
                dst.append(np.arange(ms[k], ms[k + 1]))
        old_eta = self.eta.clone() if old else None
        old_fs = self.fstate.clone() if old else None
        old_mol = {id(r): r._mol for r in old}
        if src:
            s_idx = torch.from_numpy(np.concatenate(src)).to(dev)
            d_idx = torch.from_numpy(np.concatenate(dst)).to(dev)
            for t in (self.pos, self.v, self.a):
                vals = t.index_select(0, s_idx)
                t.index_copy_(0, d_idx, vals)
        # new requests: upload positions, masses, velocities
        if new:
            xs = np.concatenate([r.xyz for r in new]).astype(np.float64)
            vs = np.zeros_like(xs)
            idx = []
            o = 0
            slot = {id(r): k for k, r in enumerate(reqs)}
            for r in new:
                k = slot[id(r)]
                n_ = len(r.Z)
                if r.kind == "nvt":
                    vs[o:o + n_] = maxwell_boltzmann(r.Z, self.T, r.seed, atomic_masses(r.Z))
                idx.append(np.arange(ms[k], ms[k + 1]))
                o += n_
            d_idx = torch.from_numpy(np.concatenate(idx)).to(dev)
            self.pos.index_copy_(0, d_idx, torch.from_numpy(xs).to(dev))
            self.v.index_copy_(0, d_idx, torch.from_numpy(vs).to(dev))
            self.a.index_copy_(0, d_idx, torch.zeros_like(torch.from_numpy(xs)).to(dev))
        Zall = np.concatenate([r.Z for r in reqs]) if reqs else np.zeros(0, np.int32)
        if N:
            self.m[:N].copy_(torch.from_numpy(atomic_masses(Zall)))
        # per-molecule state
        eta = np.zeros(B)
        kT = np.zeros(B)
        Q = np.ones(B)
        fs = np.zeros((B, 4))
        eta_old = old_eta.cpu().numpy() if old_eta is not None else None
        fs_old = old_fs.cpu().numpy() if old_fs is not None else None
        fdt0 = self.fire[0]
        for k, r in enumerate(reqs):
            n_ = len(r.Z)
            if r.kind == "nvt":
                kT[k] = IDEALGASR * 1e-10 * self.T
                tau = 20.0 * self.dt * n_
                Q[k] = kT[k] * tau * tau
                if id(r) in old_mol:
                    eta[k] = eta_old[old_mol[id(r)]]
            elif r.kind == "opt":
                fs[k] = fs_old[old_mol[id(r)]] if id(r) in old_mol else (fdt0, 0.1, 0.0, 1.0)
        if B:
            self.eta[:B].copy_(torch.from_numpy(eta))
            self.kT[:B].copy_(torch.from_numpy(kT))
            self.Q[:B].copy_(torch.from_numpy(Q))
            self.fstate[:B].copy_(torch.from_numpy(fs))
        for k, r in enumerate(reqs):
            r._off = int(ms[k])
            r._mol = k
        self.active = reqs
        self.ms = ms
        self.groups = [sum(1 for r in reqs if r.kind == g) for g in ("nvt", "opt", "sp")]
        if reqs:
            self.eng.set_topology(self.ws, Zall, ms)

    # ----------------------------------------------------------------- step
    def _step(self):
        e, ws, L = self.eng, self.ws, self.eng.lib
        st = e._stream()
        bn, bo, _ = self.groups
        ms = self.ms
        n_nvt = int(ms[bn])
        n_opt0, n_opt1 = int(ms[bn]), int(ms[bn + bo])
        if bn:
            L.tm_nh_drift(n_nvt, self.dt, _P(ws.mol), _P(self.eta), _P(self.m), _P(self.pos),
                          _P(self.v), _P(self.a), _P(self.vh), _P(self.ke2), st)
        e.run(ws, self.pos, check=False)
        if bn:
            L.tm_mol_sum_f64(bn, _P(ws.mol_start), 2, _P(self.ke2), 2, _P(self.ke2sum), 0, _P(ws.red),
                             n_nvt, st)
            L.tm_nh_eta(bn, _P(ws.mol_start), self.dt, _P(self.kT), _P(self.Q), _P(self.ke2sum),
                        _P(self.eta), st)
            L.tm_nh_kick(n_nvt, self.dt, _P(ws.mol), _P(self.eta), _P(ws.force), _P(self.m),
                         _P(self.vh), _P(self.v), _P(self.a), _P(self.ke), st)
        if bo:
            n = n_opt1 - n_opt0
            off3 = lambda t: _lib.ctypes.c_void_p(t.data_ptr() + 8 * 3 * n_opt0)
            L.tm_fire_dots(n, off3(ws.force), off3(self.v), off3(self.dots), st)
            L.tm_mol_sum_f64(bo, _lib.ctypes.c_void_p(ws.mol_start.data_ptr() + 4 * bn), 3,
                             _P(self.dots), 3, _lib.ctypes.c_void_p(self.dsum.data_ptr() + 8 * 3 * bn),
                             0, None, n, st)
            fdt0, _, fmaxstep = self.fire
            L.tm_fire_update(n, bo, _lib.ctypes.c_void_p(ws.mol.data_ptr() + 4 * n_opt0), bn,
                             _lib.ctypes.c_void_p(self.dsum.data_ptr() + 8 * 3 * bn),
                             _lib.ctypes.c_void_p(self.fstate.data_ptr() + 8 * 4 * bn),
                             _lib.ctypes.c_void_p(self.mix.data_ptr() + 8 * 3 * bn), off3(ws.force),
                             _lib.ctypes.c_void_p(self.m.data_ptr() + 8 * n_opt0), off3(self.v),
                             off3(self.pos), 10.0 * fdt0, 5, 1.1, 0.5, 0.1, 0.99, fmaxstep, st)

    # ----------------------------------------------------------------- loop
    def run(self, requests, realtime=True, max_iters=10 ** 7, on_iter=None):
        """Serve `requests` (sorted by arrival).  With realtime, arrivals follow
        the wall clock; otherwise every request is available at once."""
        pending = sorted(requests, key=lambda r: r.t_arrival)
        for r in pending:
            if len(r.Z) > self.max_atoms:
                raise ValueError(f"request {r.rid}: {len(r.Z)} atoms exceed the server capacity "
                                 f"of {self.max_atoms} atoms")
            if r.kind not in KIND_ORDER:
                raise ValueError(f"request {r.rid}: unknown kind {r.kind!r}")
        pi = 0
        done = []
        t0 = time.perf_counter()
        it = 0
        fdt0, fmax, _ = self.fire
        fthr2 = (fmax * JOULEPERHARTREE) ** 2
        self.active = []
        self._dirty = False
        self.atom_steps = 0
        while (pi < len(pending) or self.active) and it < max_iters:
            now = time.perf_counter() - t0
            new = []
            natoms = sum(len(r.Z) for r in self.active)
            nmol = len(self.active)
            while pi < len(pending) and (not realtime or pending[pi].t_arrival <= now):
                r = pending[pi]
                if natoms + len(r.Z) > self.max_atoms or nmol + 1 > self.max_mols:
                    break   # admission by capacity (HBM budget); retry next iteration
                r.t_admit = now
                new.append(r)
                natoms += len(r.Z)
                nmol += 1
                pi += 1
            if not self.active and not new:
                if pi < len(pending):
                    wait = pending[pi].t_arrival - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.001))
                continue
            if new or self._dirty:
                self._repack(new)
                self._dirty = False
            self._step()
            it += 1
            ws = self.ws
            bn, bo, _ = self.groups
            B = len(self.active)
            # one synchronising copy: batch flags, per-molecule energies, FIRE |F|^2
            parts = [ws.flags[:4].double(), self.eng.energy_tensor(ws)]
            if bo:
                parts.append(self.dsum[bn:bn + bo, 1])
            stat = torch.cat(parts).cpu().numpy()
            flags, E, conv = stat[:4], stat[4:4 + B], stat[4 + B:]
            # failures stay with their molecule: a row overflow (a collapsed
            # geometry) or a non-finite energy fails that request only, since
            # rows, mirror slots and reductions never cross molecules
            bad = ~np.isfinite(E)
            if bo:
                bad[bn:bn + bo] |= ~np.isfinite(conv)
            if flags[0] or flags[1]:
                nc = ws.ncnt[:ws.N].cpu().numpy()
                ac = ws.acnt[:ws.N].cpu().numpy()
                full = (nc >= self.eng.ldr) | (ac >= self.eng.lda)
                bad[np.unique(ws.mol_np[full])] = True
            self.atom_steps += ws.N
            keep, fin = [], []
            for k, r in enumerate(self.active):
                r.steps_done += 1
                if bad[k]:
                    r.error = ("neighbour row overflow (collapsed geometry)" if flags[0] or flags[1]
                               else "non-finite energy or forces")
                    fin.append((k, r))
                    continue
                if r.kind == "sp":
                    finished = True
                elif r.kind == "nvt":
                    finished = r.steps_done >= r.nsteps
                else:
                    c = conv[k - bn] / len(r.Z) < fthr2
                    finished = c or r.steps_done >= r.nsteps
                    r.converged = bool(c)
                (fin if finished else keep).append((k, r))
            if fin:
                # sp results sit at the end of the batch, nvt/opt at the front
                n_dyn = int(self.ms[bn + bo])
                F = ws.force[n_dyn:ws.N].cpu().numpy() if any(
                    r.kind == "sp" and r.error is None for _, r in fin) else None
                X = self.pos[:n_dyn].cpu().numpy() if any(
                    r.kind != "sp" and r.error is None for _, r in fin) else None
                t_done = time.perf_counter() - t0
                for k, r in fin:
                    a, b_ = int(self.ms[k]), int(self.ms[k + 1])
                    if r.error is None:
                        r.energy = float(E[k])
                        if r.kind == "sp":
                            r.forces = F[a - n_dyn:b_ - n_dyn].copy()
                        else:
                            r.positions = X[a:b_].copy()
                    r.t_done = t_done
                    r._off = -1
                    done.append(r)
            # finished requests leave: survivors keep their device offsets and
            # slots until the next re-pack
            if fin:
                self._dirty = True
            self.active = [r for _, r in keep]
            if on_iter is not None:
                on_iter(self, it)
        self.wall = time.perf_counter() - t0
        self.iters = it
        return done


def latency_stats(done):
    """Arrival-to-result latency percentiles of the completed requests."""
    lat = np.array([r.t_done - r.t_arrival for r in done if r.error is None]) * 1e3
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean())}


def atom_steps(done):
    """Atom-steps of the completed requests (failed requests do not count)."""
    return int(sum(len(r.Z) * r.steps_done for r in done if r.error is None))
