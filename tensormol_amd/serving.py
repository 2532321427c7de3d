"""Continuous-batching E+F server for a stream of molecule requests.

Request kinds: single-point energy + forces ("sp"), Nose-Hoover NVT segments
("nvt", per-molecule thermostat, SimpleMD.py:90-129 semantics) and FIRE
optimisations ("opt").  The reference has no batching: its only request loop
is the one-at-a-time i-PI client (TMIPIinterface.py:19-61), and a new molecule
size rebuilds its TensorFlow graph (TFMolInstanceDirect.py:5694-5699).

Every iteration of Server.run:
  1. forms the next batch from the requests of the batch on the GPU that will
     still be running after it (single points leave after one step, NVT and
     FIRE at their step limits) plus the requests that have arrived, while
     the packed batch stays within the atom capacity derived from an HBM
     budget;
  2. re-packs it into kind-contiguous groups [nvt | opt | sp] (state of
     continuing requests moves with on-device gathers queued behind the
     running step; new requests' coordinates are uploaded once, through
     pinned staging, so the host never waits for the GPU here);
  3. queues one E+F launch sequence over all atoms, the NVT integrator
     kernels on the nvt range and the FIRE kernels on the opt range, and
     copies of the step's status and results into pinned memory;
  4. waits for the previous step's status and retires its finished requests
     (single points after one evaluation, NVT after their step count, FIRE
     once the molecule's RMS atomic force is below fire_fmax Ha/A, or at the
     step limit), copying their results out and recording the latency.
A request found finished in 4 that step 1 did not foresee (a converged FIRE
run, a failure) is already in the next batch: it rides along once more and
that step's output for it is discarded, so its result is the state at the
step that finished it.
A request whose neighbour rows reach the engine's capacity (a collapsed
geometry) is not failed: it is run again from its inputs by a second server
whose engine has wider rows (256 / 128 by default), on a stream of its own,
one of its iterations after each of the batch loop's; its answer is that run's.
The rows of both engines hold the same neighbours in the same order, so a
request that fits both gives the same result on either.
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
    rescued: bool = False   # answered by the wide-row server after a row overflow
    # server bookkeeping: atom offset and molecule slot in the current packed batch
    _off: int = field(default=-1, repr=False)
    _mol: int = field(default=-1, repr=False)
    _retired: bool = field(default=False, repr=False)


def make_stream(n=4096, rate=2000.0, lo=8, hi=256, seed=3, kinds=STREAM_KINDS, workers=None):
    """The seeded request stream of workloads.request_stream as Request objects."""
    return [Request(r, Z, x, kind, steps, t, seed=sd)
            for r, (Z, x, kind, steps, t, sd) in enumerate(
                request_stream(n, rate, lo, hi, seed, kinds, workers))]


class Server:
    """Continuous batching over one engine workspace of fixed capacity."""

    def __init__(self, engine, max_atoms=65536, max_mols=8192, hbm_budget_bytes=None, T=300.0,
                 dt=0.5, fire_dt0=1.0, fire_fmax=5e-3, fire_maxstep=0.1, repack_every=1,
                 graphs=False, rescue_rows=(256, 128)):
        """repack_every: with K > 1, a batch whose membership changes only by
        departures and arrivals is repacked at most every K iterations; in
        between, the previous layout is relaunched with departed requests as
        discarded ghosts and arrivals wait (at most K - 1 iterations).  A repack
        is forced when ghosts exceed a quarter of the batch's atoms.
        graphs: a relaunched layout's step (every kernel of the evaluation and
        the integrators) is captured into a HIP graph on its first relaunch
        and replayed for the following ones, so an unchanged batch costs one
        graph launch instead of one launch per kernel.
        rescue_rows: (ldr, lda) of the engine that re-runs requests whose rows
        overflow this engine's (None: such requests fail)."""
        self.eng = engine
        self.rescue_rows = None
        if rescue_rows is not None and (rescue_rows[0] > engine.ldr or rescue_rows[1] > engine.lda):
            self.rescue_rows = (int(rescue_rows[0]), int(rescue_rows[1]))
        self._rescue = None        # the wide-row server, made on the first overflow
        self.rescued = 0           # requests answered by it in the last run
        self.repack_every = max(1, int(repack_every))
        self.use_graphs = bool(graphs)
        self._layout = 0           # incremented by every repack
        self._graph = None
        self._graph_layout = -1
        self.graph_captures = 0
        self.graph_replays = 0
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
        self.natoms = 0
        from .staging import PackedOutputs, PinnedStager
        self.stager = PinnedStager(dev, slot_bytes=96 * n + 96 * b + (1 << 16))
        # two sets of result buffers: the step on the GPU and the one being read.
        # Per step: energies (Ebp + Eext, 24 B a molecule), FIRE |F|^2 rows
        # (24 B), forces or positions (24 B an atom) and two row counts (8 B)
        pack = 48 * b + 32 * n + 256
        self.outs = [PackedOutputs(dev, pack), PackedOutputs(dev, pack)]

    def bytes_per_atom(self):
        """Device bytes of one atom's share of the workspace and server state."""
        e = self.eng
        ws = 4 * e.D * 3 + e.ldr * (4 + 4 + 16) + e.lda * 4 + 4 * 2 * (e.nele + 1) + 200
        return ws + 8 * (3 * 4 + 2 + 2 + 3)

    # ---------------------------------------------------------------- packing
    @staticmethod
    def _ranges(starts, lens):
        """Concatenation of arange(starts[k], starts[k] + lens[k]) over k."""
        tot = int(lens.sum())
        if tot == 0:
            return np.zeros(0, dtype=np.int64)
        rel = np.arange(tot, dtype=np.int64) - np.repeat(np.cumsum(lens) - lens, lens)
        return np.repeat(starts, lens) + rel

    def _repack(self, keep, new):
        """Order continuing + new requests as [nvt | opt | sp] and move device
        state.  Every host array of the batch is staged first and lands with one
        host->device copy; the device gathers and scatters are queued after it
        (all behind the running step)."""
        reqs = sorted(keep + new, key=lambda r: (KIND_ORDER[r.kind], r.rid))
        up = self.stager.put
        self.stager.begin()
        B = len(reqs)
        sizes = np.fromiter((len(r.Z) for r in reqs), dtype=np.int64, count=B)
        ms = np.zeros(B + 1, dtype=np.int64)
        np.cumsum(sizes, out=ms[1:])
        N = int(ms[-1])
        kinds = np.fromiter((KIND_ORDER[r.kind] for r in reqs), dtype=np.int64, count=B)
        old_off = np.fromiter((r._off for r in reqs), dtype=np.int64, count=B)
        old_slot = np.fromiter((r._mol for r in reqs), dtype=np.int64, count=B)
        cont = old_off >= 0
        # ---- stage (host only)
        has_cont = bool(cont.any())
        if has_cont:   # continuing requests: their atom and molecule state is gathered on the device
            s_idx = up(self._ranges(old_off[cont], sizes[cont]))
            d_idx = up(self._ranges(ms[:-1][cont], sizes[cont]))
            so = up(old_slot[cont])
            sn = up(np.nonzero(cont)[0])
        newk = np.nonzero(~cont)[0]
        if len(newk):
            xs = np.concatenate([reqs[k].xyz for k in newk]).astype(np.float64)
            vs = np.zeros_like(xs)
            o = 0
            for k in newk:
                r = reqs[k]
                if r.kind == "nvt":
                    vs[o:o + len(r.Z)] = maxwell_boltzmann(r.Z, self.T, r.seed, atomic_masses(r.Z))
                o += len(r.Z)
            d_new = up(self._ranges(ms[:-1][newk], sizes[newk]))
            xs_d, vs_d = up(xs), up(vs)
        Zall = np.concatenate([r.Z for r in reqs]) if reqs else np.zeros(0, np.int32)
        # per-molecule state: thermostat constants from kind and size (SimpleMD.py:109-125),
        # fresh FIRE state for new optimisations, carried state for continuing requests
        kT = np.where(kinds == 0, IDEALGASR * 1e-10 * self.T, 0.0)
        tau = 20.0 * self.dt * sizes
        Q = np.where(kinds == 0, kT * tau * tau, 1.0)
        fs = np.zeros((B, 4))
        fs[kinds == 1] = (self.fire[0], 0.1, 0.0, 1.0)
        if N:
            self.m = up(atomic_masses(Zall))
        if B:
            self.kT = up(kT)
            self.Q = up(Q)
            fs_d = up(fs)
        if reqs:
            self.eng.set_topology(self.ws, Zall, ms, stager=self.stager)
        self.stager.flush()
        # ---- device work, queued after the upload
        if has_cont:
            for t in (self.pos, self.v, self.a):
                t.index_copy_(0, d_idx, t.index_select(0, s_idx))
            eta_c = self.eta.index_select(0, so)
            fs_c = self.fstate.index_select(0, so)
        if len(newk):
            self.pos.index_copy_(0, d_new, xs_d)
            self.v.index_copy_(0, d_new, vs_d)
            self.a.index_fill_(0, d_new, 0.0)
        if B:
            self.eta[:B].zero_()
            self.fstate[:B].copy_(fs_d)
            if has_cont:
                self.eta.index_copy_(0, sn, eta_c)
                self.fstate.index_copy_(0, sn, fs_c)
        for k, r in enumerate(reqs):
            r._off = int(ms[k])
            r._mol = k
        self.active = reqs
        self.ms = ms
        self.natoms = N
        self.groups = [int(c) for c in np.bincount(kinds, minlength=3)]
        self._layout += 1
        self.stager.end()

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
    def _capture(self):
        """Capture one step of the current layout into a HIP graph (the
        kernels only: the result copies stay outside, since they alternate
        between two pinned buffers)."""
        cur = torch.cuda.current_stream(self.eng.device)
        side = getattr(self, "_cap_stream", None)
        if side is None:
            side = self._cap_stream = torch.cuda.Stream(device=self.eng.device)
        side.wait_stream(cur)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(side):
            g.capture_begin()
            try:
                self._step()
            finally:
                g.capture_end()
        cur.wait_stream(side)
        self._graph, self._graph_layout = g, self._layout
        self.graph_captures += 1

    def _launch(self, outs, relaunch=False):
        """Queue one step of the current batch and the copies of its status
        (flags, per-molecule energies, FIRE |F|^2), the single points' forces,
        the nvt/opt positions and the row counts into pinned memory.  A
        relaunched layout runs as a HIP graph when graphs are on."""
        ws = self.ws
        bn, bo, _ = self.groups
        if self.use_graphs and relaunch:
            if self._graph_layout != self._layout:
                self._capture()
            self._graph.replay()
            self.graph_replays += 1
        else:
            self._step()
        B, N = ws.B, ws.N
        n_dyn = int(self.ms[bn + bo])
        outs.begin()
        outs.add("Ebp", ws.Ebp[:B])
        outs.add("Eext", ws.Eext[:B])
        outs.add("dsum", self.dsum[bn:bn + bo])
        outs.add("force", ws.force[n_dyn:N])
        outs.add("pos", self.pos[:n_dyn])
        outs.add("ncnt", ws.ncnt[:N])
        outs.add("acnt", ws.acnt[:N])
        outs.mark()
        return dict(reqs=self.active, ms=self.ms, groups=tuple(self.groups), N=ws.N,
                    mol_np=ws.mol_np, outs=outs)

    @staticmethod
    def _leaves_after(r):
        """Whether r leaves after its next step whatever that step computes."""
        return r.kind == "sp" or r.steps_done + 1 >= r.nsteps

    def _finish(self, batch, t0, fthr2, done):
        """Wait for a launched step and retire its finished requests."""
        outs = batch["outs"]
        outs.wait()
        reqs, ms, (bn, bo, _) = batch["reqs"], batch["ms"], batch["groups"]
        B, N = len(reqs), batch["N"]
        # the engine's molecule total (Engine.energy_tensor), same order of adds
        Eext = outs.host("Eext")
        E = outs.host("Ebp") + Eext[:, 0] + Eext[:, 1]
        conv = outs.host("dsum")[:, 1]
        # failures stay with their molecule, since rows, mirror slots and
        # reductions never cross molecules.  A molecule fails when one of its
        # rows reaches capacity (ldr - 1 and lda - 1 neighbours are the most a
        # request may have: a collapsed geometry), judged per molecule so the
        # verdict does not depend on the batch, or when its energy is not finite
        full = (outs.host("ncnt") >= self.eng.ldr) | (outs.host("acnt") >= self.eng.lda)
        row_full = np.zeros(B, dtype=bool)
        row_full[np.unique(batch["mol_np"][full])] = True
        bad = row_full | ~np.isfinite(E)
        if bo:
            bad[bn:bn + bo] |= ~np.isfinite(conv)
        n_dyn = int(ms[bn + bo])
        t_done = time.perf_counter() - t0
        for k, r in enumerate(reqs):
            if r._retired:      # finished at the previous step; this output is discarded
                continue
            r.steps_done += 1
            if bad[k] and row_full[k] and self.rescue_rows is not None:
                if self._rescue is None:
                    self._rescue = _RowRescue(self)
                self._rescue.submit(r, t0)   # answered by the wide-row server
                r._retired = True
                continue
            if bad[k]:
                r.error = ("neighbour row at capacity (collapsed geometry)" if row_full[k]
                           else "non-finite energy or forces")
                finished = True
            elif r.kind == "sp":
                finished = True
            elif r.kind == "nvt":
                finished = r.steps_done >= r.nsteps
            else:
                c = conv[k - bn] / len(r.Z) < fthr2
                finished = c or r.steps_done >= r.nsteps
                r.converged = bool(c)
            if not finished:
                continue
            a, b_ = int(ms[k]), int(ms[k + 1])
            if r.error is None:
                r.energy = float(E[k])
                if r.kind == "sp":
                    r.forces = outs.host("force")[a - n_dyn:b_ - n_dyn].copy()
                else:
                    r.positions = outs.host("pos")[a:b_].copy()
            r.t_done = t_done
            r._retired = True
            done.append(r)

    def run(self, requests, realtime=True, max_iters=10 ** 7, on_iter=None):
        """Serve `requests`.  With realtime, arrivals follow the wall clock;
        otherwise every request is available at once."""
        gen = self.iterate(requests, realtime, max_iters, on_iter)
        while True:
            try:
                next(gen)
            except StopIteration as stop:
                return stop.value

    def iterate(self, requests, realtime=True, max_iters=10 ** 7, on_iter=None):
        """run() as a generator: yields after every iteration of the batch loop
        and returns the finished requests."""
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
        _, fmax, _ = self.fire
        fthr2 = (fmax * JOULEPERHARTREE) ** 2
        self.active = []
        self.natoms = 0
        self.atom_steps = 0
        inflight = None
        since = 0   # iterations since the last repack
        while it < max_iters:
            # continuing requests: those of the launched step that outlive it
            if inflight is not None:
                keep = [r for r in inflight["reqs"] if not r._retired and not self._leaves_after(r)]
            else:
                keep = []
            natoms = sum(len(r.Z) for r in keep)
            nmol = len(keep)
            now = time.perf_counter() - t0
            # relaunch the previous layout (departed requests as ghosts, arrivals
            # waiting) unless a repack is due
            relaunch = False
            if self.repack_every > 1 and inflight is not None and keep and since + 1 < self.repack_every:
                relaunch = 4 * natoms >= 3 * inflight["N"]
            new = []
            while not relaunch and pi < len(pending) and (not realtime or pending[pi].t_arrival <= now):
                r = pending[pi]
                if natoms + len(r.Z) > self.max_atoms or nmol + 1 > self.max_mols:
                    break   # admission by capacity (HBM budget); retried next iteration
                r.t_admit = now
                r._off = r._mol = -1
                r._retired = False
                new.append(r)
                natoms += len(r.Z)
                nmol += 1
                pi += 1
            nxt = None
            if keep or new:
                if relaunch:
                    # requests whose last step is the one on the GPU are retired by
                    # its _finish, before this launch's results are read
                    since += 1
                elif new or inflight is None or len(keep) != len(inflight["reqs"]):
                    self._repack(keep, new)
                    since = 0
                self.outs.reverse()
                nxt = self._launch(self.outs[0], relaunch=relaunch)
                it += 1
                self.atom_steps += nxt["N"]
            if inflight is not None:
                self._finish(inflight, t0, fthr2, done)
                if on_iter is not None:
                    on_iter(self, it)
            if self._rescue is not None:
                self._rescue.step()
                done.extend(self._rescue.take())
            inflight = nxt
            if inflight is None:
                if pi >= len(pending):
                    break
                wait = pending[pi].t_arrival - (time.perf_counter() - t0)
                if realtime and wait > 0:
                    time.sleep(min(wait, 0.001))
            yield it
        if inflight is not None:
            self._finish(inflight, t0, fthr2, done)
        self.rescued = 0
        if self._rescue is not None:
            done.extend(self._rescue.close())
            self.rescued = self._rescue.count
            self._rescue = None
        self.active = []
        self.wall = time.perf_counter() - t0
        self.iters = it
        return done


class _RowRescue:
    """Requests whose neighbour rows overflowed, run again from their inputs on
    a server whose engine has wider rows, on a stream of its own.  Its batch
    loop advances one iteration per call of step(); the original request is
    answered with that run's result (its latency runs to that moment)."""

    def __init__(self, server):
        from .engine import BPEEEngine
        e = server.eng
        ldr, lda = server.rescue_rows
        eng = BPEEEngine(e.model, device=e.device, precision=e.precision, ldr=ldr, lda=lda,
                         check_overflow=e.check_overflow)
        fdt0, fmax, fmaxstep = server.fire
        # its batches change only by departures: relaunched layouts replay as graphs
        self.server = Server(eng, max_atoms=server.max_atoms, max_mols=min(server.max_mols, 1024),
                             T=server.T, dt=server.dt, fire_dt0=fdt0, fire_fmax=fmax,
                             fire_maxstep=fmaxstep, rescue_rows=None, repack_every=4,
                             graphs=True)
        self.stream = torch.cuda.Stream(device=eng.device)
        self.queue = []     # (original, copy, t0) waiting for the next batch
        self.batch = None   # those of the batch being run
        self.gen = None
        self.out = []
        self.count = 0

    def submit(self, orig, t0):
        copy = Request(orig.rid, orig.Z, orig.xyz, orig.kind, orig.nsteps, 0.0, seed=orig.seed)
        self.queue.append((orig, copy, t0))
        self.count += 1

    def step(self):
        """One iteration of the wide-row batch loop (a new batch of the queued
        requests when none is running)."""
        if self.gen is None:
            if not self.queue:
                return
            self.batch, self.queue = self.queue, []
            self.gen = self.server.iterate([c for _, c, _ in self.batch], realtime=False)
        with torch.cuda.stream(self.stream):
            try:
                next(self.gen)
                return
            except StopIteration:
                pass
        now = time.perf_counter()
        for orig, c, t0 in self.batch:
            orig.energy, orig.forces, orig.positions = c.energy, c.forces, c.positions
            orig.steps_done, orig.converged, orig.error = c.steps_done, c.converged, c.error
            orig.rescued = True
            orig.t_done = now - t0
            self.out.append(orig)
        self.gen = self.batch = None

    def take(self):
        """Requests answered since the last call."""
        out, self.out = self.out, []
        return out

    def close(self):
        """Runs the remaining requests to the end; returns those not yet taken."""
        while self.gen is not None or self.queue:
            self.step()
        torch.cuda.current_stream(self.server.eng.device).wait_stream(self.stream)
        return self.take()


def latency_stats(done):
    """Arrival-to-answer latency percentiles over every finished request, the
    failed ones included (a failure is an answer too, at the time it is
    reported), and the failure count."""
    lat = np.array([r.t_done - r.t_arrival for r in done]) * 1e3
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "failed": int(sum(r.error is not None for r in done))}


def atom_steps(done):
    """Atom-steps of the completed requests (failed requests do not count)."""
    return int(sum(len(r.Z) * r.steps_done for r in done if r.error is None))
