"""Benchmark: NVE molecular dynamics of the periodic 4096-water box (12,288 atoms)
on one MI355X, BP+EE energy and forces every step.

Workload: rigid waters on a jittered 16^3 lattice in a 49.7 A cube (seed 1),
minimum-image cell lists, 2-element descriptor of length 256, energy and
charge nets 256-512-512-512-1 softplus with seeded Glorot weights, DSF alpha
0.2/A with a 15 A cutoff, Maxwell-Boltzmann velocities at 300 K, Velocity
Verlet at dt = 0.5 fs.  Descriptors FP32, MLP GEMMs BF16 on MFMA with FP32
accumulation, forces by the hand-written adjoint, FP64 positions, velocities
and energy reductions.

A step is one Velocity-Verlet step: drift, neighbour rebuild, full E+F, kick.
value = atoms * timed steps / wall time of the timed steps (inputs resident in
HBM).  The JSON line also carries the force RMS error of frame 0 against the
FP64 CPU oracle, the plain-FP32 path's rate on the same box, the roofline of
the MLP kernel and the CPU baseline (the oracle, timed on this host).

Serving leg ("serving", and the top-level p50_ms / p99_ms request latencies):
a seeded stream of 4096 C/H/N/O chain molecules (8-256 atoms, log-uniform;
50% single points, 30% 50-step Nose-Hoover NVT, 20% 100-step FIRE) arriving
as a Poisson process at 2000 requests/s, served by the continuous-batching
server with the CHNO model (768-512-512-512-1 nets).  Reported: latency
percentiles (arrival to result), atom-steps/s of the realtime stream and of
the saturated server (all requests queued at t = 0), and the plain FP32 path
serving the first requests of the stream one molecule per launch.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MFMA operand type of each MLP precision mode (FP32 accumulation throughout)
MFMA_DTYPE = {"bf16": "bf16", "bf16x2": "bf16", "bf16x3f": "bf16", "bf16x3": "bf16",
              "fp16": "fp16", "fp16x2": "fp16", "fp32": "fp32"}
MLP_PRECISION_NOTE = {
    "bf16x3f": "BF16 MFMA, FP32 accumulation; forward GEMMs on hi/lo BF16 splits of weights and "
               "activations (three MFMAs per product), input-gradient GEMMs plain BF16",
    "bf16": "BF16 MFMA, FP32 accumulation, operands rounded once",
}
FP32_VALU_PEAK_TFLOPS = 157.3
COUL_FLOPS_PER_PAIR = 50     # FP32 operations per unordered pair within 15 A (erfc/exp included)
MLP_PEAK_TFLOPS = {"bf16": 2500.0, "fp16": 2500.0, "bf16x3": 2500.0, "bf16x2": 2500.0, "fp16x2": 2500.0,
                   "bf16x3f": 2500.0, "fp32": 157.3}


def mlp_flops_per_atom(D, H):
    """Two nets x (forward + input gradient): 4 * 2 * (D*H + 2*H*H + H)."""
    return 4 * 2 * (D * H + 2 * H * H + H)


def build_system(args):
    from tensormol_amd import workloads as W
    from tensormol_amd.model import BPEEModel
    Z, x, box = W.water_box(nside=args.nside, edge=args.edge, seed=1)
    model = BPEEModel([1, 8], dsf_alpha=0.2, seed=0)
    return Z, x, box, model


def run_md(eng, ws, Z, x, box, steps, warmup, dt, graph=True, mlp_samples=20, steps_per_graph=10):
    """NVE Velocity Verlet on the device.  Returns wall time of the timed steps,
    per-step times (HIP events between graph replays, divided by the steps a
    replay holds), live launch times of the MLP stage and of the Coulomb sweep
    (HIP events around the launches, on the stream they run on) and the total
    energy before and after the timed steps.  With graph, `steps_per_graph` consecutive steps (the largest
    divisor of `steps` not above it) are captured into one HIP graph, so the
    replay's launch cost is paid once per that many steps."""
    import torch
    from tensormol_amd import _lib, params
    from tensormol_amd import workloads as W
    dev = eng.device
    m = torch.tensor(params.atomic_masses(Z), device=dev)
    v0 = W.maxwell_boltzmann(Z, 300.0, 1, params.atomic_masses(Z))
    pos = torch.tensor(x, dtype=torch.float64, device=dev)
    v = torch.tensor(v0, device=dev)
    a = torch.zeros_like(v)
    ke = torch.zeros(ws.N, dtype=torch.float64, device=dev)
    L = eng.lib
    P = _lib.ptr
    side = torch.cuda.Stream(device=dev)
    side.wait_stream(torch.cuda.current_stream(dev))
    mlp_ms = []
    coul_ms = []
    with torch.cuda.stream(side):
        st = eng._stream()
        # initial forces -> accelerations (a = F/m before the first drift), and
        # the first drift: every step is then evaluate + kick + the next drift
        # (the kick and drift run inside the force gather's launch), the same
        # operations as drift + evaluate + kick
        eng.run(ws, pos)
        L.tm_vv_kick(ws.N, 0.0, P(ws.force), P(m), P(v), P(a), P(ke), st)
        L.tm_vv_drift(ws.N, dt, P(v), P(a), P(pos), st)

        def step(record=False, clear_flags=True):
            eng.build_lists(ws, pos, clear_flags=clear_flags)
            eng.descriptors(ws, pos)
            if record:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(side)
            eng.networks(ws)
            timer = None
            if record:
                e1.record(side)
                mlp_ms.append((e0, e1))
                c0 = torch.cuda.Event(enable_timing=True)
                c1 = torch.cuda.Event(enable_timing=True)
                coul_ms.append((c0, c1))
                timer = (lambda: c0.record(side), lambda: c1.record(side))
            eng.electrostatics(ws, pos, timer=timer)
            eng.adjoint(ws, pos, vv=(dt, m, v, a, ke, pos))
            eng.reductions(ws)

        for _ in range(warmup):
            step()
        # live launch time of the MLP kernel, on the stream it is launched on
        for _ in range(mlp_samples):
            step(record=True)
        eng.check_flags(ws)
        g = None
        if graph:
            g = torch.cuda.CUDAGraph()
            side.synchronize()
    spg = 1
    if graph:
        spg = max(1, min(int(steps_per_graph), steps))
        while steps % spg:
            spg -= 1
        # the status flags are cleared by the first step of a replay and
        # accumulate over its steps, so the check after the replays covers all
        with torch.cuda.graph(g, stream=side):
            for k in range(spg):
                step(clear_flags=k == 0)
        replay = g.replay
    else:
        replay = step
    side.synchronize()
    with torch.cuda.stream(side):
        e_start = (eng.energy_tensor(ws).item(), ke.sum().item())
        step_ev = []
        first = torch.cuda.Event(enable_timing=True)
        side.synchronize()
        t0 = time.perf_counter()
        first.record(side)
        for _ in range(steps // spg):
            replay()
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(side)
            step_ev.append(ev)
        side.synchronize()
        t1 = time.perf_counter()
        eng.check_flags(ws)
        e_end = (eng.energy_tensor(ws).item(), ke.sum().item())
    per_step = []
    prev = first
    for ev in step_ev:
        per_step.append(prev.elapsed_time(ev) / spg)
        prev = ev
    mlp = np.array([a_.elapsed_time(b_) for a_, b_ in mlp_ms])
    coul = np.array([a_.elapsed_time(b_) for a_, b_ in coul_ms])
    return dict(wall=t1 - t0, per_step_ms=np.array(per_step), mlp_ms=mlp, coul_ms=coul,
                e_start=e_start, e_end=e_end, graph=graph, steps_per_graph=spg)


STAGES = ("build_lists", "descriptors", "networks", "electrostatics", "adjoint", "reductions")


def stage_times(eng, ws, pos, reps, stream):
    """Median GPU time (ms) of each engine stage, from HIP events recorded on
    the stream the stages run on, over `reps` eager evaluations (the median
    keeps one stalled launch from the host out of the figure)."""
    import torch
    recs = []
    with torch.cuda.stream(stream):
        for _ in range(reps):
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(STAGES) + 1)]
            ev[0].record(stream)
            for k, name in enumerate(STAGES):
                fn = getattr(eng, name)
                fn(ws) if name in ("networks", "reductions") else fn(ws, pos)
                ev[k + 1].record(stream)
            recs.append(ev)
    stream.synchronize()
    out = {}
    for k, name in enumerate(STAGES):
        out[name] = float(np.median([e[k].elapsed_time(e[k + 1]) for e in recs]))
    return out


def descriptor_bytes(N, D, P_r, P_r3):
    """Algorithmic HBM bytes of the descriptor forward pass: positions (12N),
    row offsets (4(N+1)), the 4.6 A and 3.1 A neighbour indices (4 P_r,
    4 P_r3; triples are enumerated on the fly) and the descriptor write (4 D N)."""
    return 12 * N + 4 * (N + 1) + 4 * P_r + 4 * P_r3 + 4 * D * N


def run_sphere(args):
    """Single-point E+F of the open-boundary C/H/N/O sphere (24,000 atoms)."""
    import torch
    from oracle import oracle as O
    from tensormol_amd import workloads as W
    from tensormol_amd.engine import BPEEEngine
    from tensormol_amd.model import BPEEModel
    log = lambda m: print(f"[sphere] {m}", file=sys.stderr, flush=True)
    Z, x = W.chno_sphere(natom=args.sphere_atoms, seed=2)
    N = len(Z)
    model = BPEEModel([1, 6, 7, 8], dsf_alpha=0.2, seed=0)
    eng = BPEEEngine(model, precision=args.sphere_precision, check_overflow=False)
    ms = np.array([0, N])
    ws = eng.prepare(Z, ms, None, pos0=x)
    dev = eng.device
    pos = torch.tensor(x, dtype=torch.float64, device=dev)
    side = torch.cuda.Stream(device=dev)
    side.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(side):
        for _ in range(3):
            eng.run(ws, pos)
        eng.check_flags(ws)
    side.synchronize()
    stages = stage_times(eng, ws, pos, 10, side)
    P_r = int(ws.ncnt[:N].sum())
    P_r3 = int(ws.acnt[:N].sum())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        eng.run(ws, pos, check=False)
    side.synchronize()
    with torch.cuda.stream(side):   # replay() launches on the current stream
        t0 = time.perf_counter()
        for _ in range(args.sphere_evals):
            g.replay()
        side.synchronize()
        wall = time.perf_counter() - t0
        eng.check_flags(ws)
        r = eng.results(ws)
    value = N * args.sphere_evals / wall
    log(f"{N} atoms {args.sphere_precision}: {1e3 * wall / args.sphere_evals:.3f} ms per E+F, "
        f"{value:.4g} atom-steps/s; stages {stages}")
    t0 = time.perf_counter()
    ref = O.evaluate(x, Z, ms, O.SymParams(model.eles), model.energy_nets, model.charge_nets,
                     alpha=model.alpha, dsf_alpha=model.dsf_alpha, has_vdw=model.has_vdw)
    cpu_s = time.perf_counter() - t0
    frms = float(np.sqrt(((r["gradient"] - ref["gradient"]) ** 2).mean()))
    dE = float(abs(r["Etotal"][0] - ref["Etotal"][0]))
    log(f"oracle {cpu_s:.1f} s: force rms err {frms:.3e} Ha/A, |dE| {dE:.3e} Ha")
    eng32 = BPEEEngine(model, precision="fp32", lda=128, check_overflow=False)
    ws32 = eng32.prepare(Z, ms, None, pos0=x)
    with torch.cuda.stream(side):
        eng32.run(ws32, pos)
    side.synchronize()
    g32 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g32, stream=side):
        eng32.run(ws32, pos, check=False)
    side.synchronize()
    with torch.cuda.stream(side):   # same graph replay as the timed FP16 evaluations
        t0 = time.perf_counter()
        for _ in range(args.sphere_evals):
            g32.replay()
        side.synchronize()
        v32 = N * args.sphere_evals / (time.perf_counter() - t0)
    D, H = eng.D, int(model.hidden[0])
    flops = mlp_flops_per_atom(D, H) * N
    bdesc = descriptor_bytes(N, D, P_r, P_r3)
    peak = MLP_PEAK_TFLOPS[args.sphere_precision]
    mlp_tf = flops / (stages["networks"] * 1e-3) / 1e12
    desc_gbs = bdesc / (stages["descriptors"] * 1e-3) / 1e9
    adj_gbs = (bdesc + 12 * N) / (stages["adjoint"] * 1e-3) / 1e9
    import torch as _t
    return {
        "metric": "atom-steps/s (E+F) at force RMS<1e-3 Ha/A, single point",
        "value": round(value, 1), "unit": "atom-steps/s", "n_gpus": 1,
        "evals": args.sphere_evals, "ms_per_eval": round(1e3 * wall / args.sphere_evals, 4),
        "higher_is_better": True, "dtype": args.sphere_precision, "data": "synthetic",
        "config": {"workload": f"open {N}-atom C/H/N/O sphere (H .5 C .3 N .1 O .1, 0.10/A^3, "
                               "min distance 1.0 A) single-point E+F, D 768, nets 768-512-512-512-1 x2",
                   "atoms": N, "hip_graph": True},
        "pairs_4.6A": P_r, "pairs_3.1A": P_r3,
        "force_rms_err_vs_fp64": frms, "energy_err_vs_fp64_Ha": dE,
        "plain_fp32_value": round(v32, 1), "speedup_vs_plain_fp32": round(value / v32, 2),
        "stage_ms": {k: round(v, 4) for k, v in stages.items()},
        "roofline": {
            "mlp": {"bound": "mfma", "achieved": round(mlp_tf, 2), "peak": peak, "unit": "TFLOP/s",
                    "frac": round(mlp_tf / peak, 4), "flops_per_launch": flops},
            "descriptors": {"bound": "hbm", "achieved": round(desc_gbs, 1), "peak": 8000.0,
                            "unit": "GB/s", "frac": round(desc_gbs / 8000.0, 4),
                            "bytes_per_launch": bdesc},
            "adjoint": {"bound": "hbm", "achieved": round(adj_gbs, 1), "peak": 8000.0, "unit": "GB/s",
                        "frac": round(adj_gbs / 8000.0, 4), "bytes_per_launch": bdesc + 12 * N}},
        "cpu_baseline": {"value": round(N / cpu_s, 1), "unit": "atom-steps/s",
                         "cores": _t.get_num_threads(), "kind": "port",
                         "sample": f"one FP64 E+F of the same sphere by the oracle, {cpu_s:.1f} s"},
    }


def run_train(args):
    """Force-matching training steps on 256 C/H/N/O chain molecules of 20-60
    atoms: labels from an FP64 teacher BP network (the oracle), student of the
    same 768-512-512-512-1 shape, loss MSE(E) + 0.5 MSE(F), Adam."""
    import torch
    from oracle import oracle as O
    from tensormol_amd import workloads as W
    from tensormol_amd.model import BPEEModel
    from tensormol_amd.train import ForceMatchingTrainer
    log = lambda m: print(f"[train] {m}", file=sys.stderr, flush=True)
    eles = [1, 6, 7, 8]
    rng = np.random.default_rng(4)
    sizes = rng.integers(20, 61, size=args.train_mols)
    Z, x, ms = W.pack(W.chain_set(sizes, seed=4))
    N, B = len(Z), len(ms) - 1
    sp = O.SymParams(eles)
    t0 = time.perf_counter()
    lists = O.build_lists(x, Z, ms, sp)
    teacher = BPEEModel(eles, dsf_alpha=0.2, seed=11)
    student = BPEEModel(eles, dsf_alpha=0.2, seed=12)
    xt = torch.tensor(x, requires_grad=True)
    Et = O.bp_energies(xt, Z, ms, sp, O.layers_torch(teacher.energy_nets), teacher.alpha, lists)
    (g,) = torch.autograd.grad(Et.sum(), xt)
    E_ref, F_ref = Et.detach().numpy(), -g.numpy()
    log(f"{B} molecules, {N} atoms; FP64 teacher labels in {time.perf_counter() - t0:.1f} s")
    out = {}
    dev = torch.device("cuda")
    pos = torch.tensor(x, dtype=torch.float64, device=dev)
    for dtype in ("bf16", "fp32"):
        # eager steps, then the same step captured into a HIP graph and replayed
        tr = ForceMatchingTrainer(student, mlp_dtype=dtype, lr=1e-4, capturable=True)
        tb = tr.batch(Z, ms, E_ref, F_ref)
        for _ in range(args.warmup):
            tr.step(tb, pos)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        losses = [tr.step(tb, pos) for _ in range(args.train_steps)]
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        losses = [float(v.detach()) for v in losses]
        rec = dict(eager_ms_per_step=1e3 * wall / args.train_steps,
                   eager_atoms_per_s=N * args.train_steps / wall)
        if not args.train_no_graph:
            g, Lg = tr.capture(tb, pos, warmup=3)
            for _ in range(2):
                g.replay()
            torch.cuda.synchronize()
            gl = []
            t0 = time.perf_counter()
            for _ in range(args.train_steps):
                g.replay()
                gl.append(Lg.clone())
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            losses += [float(v) for v in gl]
        rec.update(ms_per_step=1e3 * wall / args.train_steps, atoms_per_s=N * args.train_steps / wall,
                   hip_graph=not args.train_no_graph, loss_first=losses[0], loss_last=losses[-1])
        if dtype == "bf16":
            # the GEMM stage (every grouped GEMM launch of a step), timed live with
            # HIP events on the launch stream over eager steps
            tr.gemm_timing = []
            nst = 5
            for _ in range(nst):
                tr.step(tb, pos)
            torch.cuda.synchronize()
            gms = sum(a_.elapsed_time(b_) for a_, b_, _ in tr.gemm_timing) / nst
            gfl = sum(f for _, _, f in tr.gemm_timing) / nst
            rec.update(gemm_ms_per_step=gms, gemm_flops_per_step=gfl,
                       gemm_launches_per_step=len(tr.gemm_timing) // nst)
            tr.gemm_timing = None
        out[dtype] = rec
        log(f"{dtype}: {rec['ms_per_step']:.2f} ms/step ({'graph replay' if rec['hip_graph'] else 'eager'}), "
            f"eager {rec['eager_ms_per_step']:.2f} ms/step, {rec['atoms_per_s']:.4g} atoms/s, "
            f"loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    # gradient check on the first 16 molecules against the FP64 oracle's double backward
    k = 16
    Zs, xs, mss = Z[:ms[k]], x[:ms[k]], ms[:k + 1]
    w_e, w_f = np.full(k, 1.0 / k), np.full(k, 0.5 / (3 * len(Zs)))
    ls = O.build_lists(xs, Zs, mss, sp)
    lt = O.layers_torch(student.energy_nets, requires_grad=True)
    L, _, _ = O.force_matching_loss(xs, Zs, mss, sp, lt, student.alpha, E_ref[:k], F_ref[:ms[k]],
                                    w_e, w_f, ls)
    gref = torch.autograd.grad(L, [t for e in eles for (Wt, bt) in lt[e] for t in (Wt, bt)])
    rv = np.concatenate([t.numpy().ravel() for t in gref])
    errs = {}
    for dtype in ("bf16", "fp32"):
        tr = ForceMatchingTrainer(student, mlp_dtype=dtype)
        tb = tr.batch(Zs, mss, E_ref[:k], F_ref[:ms[k]], w_e, w_f)
        _, gg = tr.gradients(tb, torch.tensor(xs, dtype=torch.float64, device=dev))
        gv = np.concatenate([t.double().cpu().numpy().ravel() for t in gg])
        errs[dtype] = float(np.linalg.norm(gv - rv) / np.linalg.norm(rv))
    log(f"weight-gradient rel err vs FP64 oracle (16 molecules): {errs}")
    # CPU baseline: the oracle's FP64 loss and weight gradient (double backward,
    # neighbour lists included) of the whole batch, one step without the update
    t0 = time.perf_counter()
    lsf = O.build_lists(x, Z, ms, sp)
    ltf = O.layers_torch(student.energy_nets, requires_grad=True)
    Lf, _, _ = O.force_matching_loss(x, Z, ms, sp, ltf, student.alpha, E_ref, F_ref,
                                     np.full(B, 1.0 / B), np.full(B, 0.5 / (3 * N)), lsf)
    torch.autograd.grad(Lf, [t for e in eles for (Wt, bt) in ltf[e] for t in (Wt, bt)])
    cpu_s = time.perf_counter() - t0
    bf = out["bf16"]
    gtf = bf["gemm_flops_per_step"] / (bf["gemm_ms_per_step"] * 1e-3) / 1e12
    log(f"GEMM stage {bf['gemm_ms_per_step']:.3f} ms/step over {bf['gemm_launches_per_step']} "
        f"launches, {gtf:.1f} TFLOP/s; oracle loss + gradient of {N} atoms in {cpu_s:.1f} s")
    return {
        "metric": "force-matching training atoms/s (E+F loss, weight gradient by double backward)",
        "value": round(out["bf16"]["atoms_per_s"], 1), "unit": "atoms/s", "n_gpus": 1,
        "steps": args.train_steps, "warmup": args.warmup,
        "ms_per_step": round(out["bf16"]["ms_per_step"], 3), "higher_is_better": True,
        "dtype": "bf16 GEMM operands (rounded), fp32 products and accumulation", "data": "synthetic",
        "config": {"workload": f"{B} C/H/N/O chain molecules of 20-60 atoms ({N} atoms), "
                               "BP nets 768-512-512-512-1, loss MSE(E) + 0.5 MSE(F), Adam"},
        "eager_ms_per_step": round(out["bf16"]["eager_ms_per_step"], 3),
        "hip_graph": out["bf16"]["hip_graph"],
        "fp32_mlp": {k_: (round(v, 4) if isinstance(v, float) else v) for k_, v in out["fp32"].items()},
        "loss_first_last_bf16": [out["bf16"]["loss_first"], out["bf16"]["loss_last"]],
        "weight_grad_rel_err_vs_fp64": errs,
        "roofline": {"bound": "mfma", "kernel": "gemm_bf16_kernel",
                     "achieved": round(gtf, 2), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(gtf / 2500.0, 4), "traffic": None,
                     "flops_per_step": bf["gemm_flops_per_step"],
                     "gemm_ms_per_step": round(bf["gemm_ms_per_step"], 4),
                     "launches_per_step": bf["gemm_launches_per_step"],
                     "note": "the step's grouped GEMMs (forward, J chain, tangent, reverse, weight "
                             "gradients with their split sum): algorithmic 2MNK flops over the "
                             "launches' event-timed duration"},
        "cpu_baseline": {"value": round(N / cpu_s, 1), "unit": "atoms/s",
                         "cores": torch.get_num_threads(), "kind": "port",
                         "sample": f"one step's FP64 loss and weight gradient (double backward, "
                                   f"neighbour lists included, no update) of the whole batch "
                                   f"({N} atoms) by the oracle, {cpu_s:.1f} s"},
    }


def cpu_baseline(Z, x, box, model, threads=None):
    """The FP64 oracle on this host: one full E+F of the box (neighbour lists
    included), which is also the frame-0 reference for the force error.
    threads: torch CPU threads (default: every core the host reports)."""
    import torch
    from oracle import oracle as O
    before = torch.get_num_threads()
    torch.set_num_threads(threads or os.cpu_count() or 1)
    sp = O.SymParams(model.eles)
    try:
        t0 = time.perf_counter()
        r = O.evaluate(x, Z, np.array([0, len(Z)]), sp, model.energy_nets, model.charge_nets,
                       alpha=model.alpha, dsf_alpha=model.dsf_alpha, has_vdw=model.has_vdw,
                       boxes=box[None])
        dt = time.perf_counter() - t0
    finally:
        torch.set_num_threads(before)
    return r, dt


def cpu_model_name():
    """The host CPU's model name (as lscpu reports it), from /proc/cpuinfo."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def coulomb_pairs(x, box, rc=15.0):
    """Unordered minimum-image pairs closer than rc (the 15 A EE list size)."""
    from scipy.spatial import cKDTree
    L = np.asarray(box, dtype=np.float64)
    tree = cKDTree(np.mod(x, L), boxsize=L)
    return (int(tree.count_neighbors(tree, rc)) - len(x)) // 2


def serving_fp32_baselines(args, model, mk, stream, k, log):
    """Plain FP32 serving baselines: the continuous-batching server on the FP32
    path (saturated) and one molecule per launch on the first k requests;
    returns both rates in atom-steps/s."""
    import torch
    from tensormol_amd.engine import BPEEEngine
    from tensormol_amd.serving import Server, atom_steps
    # plain FP32 path: continuous batching (separates the batching and MFMA gains) ...
    eng32 = BPEEEngine(model, precision="fp32", lda=128, check_overflow=False)
    srv32 = Server(eng32, max_atoms=args.serve_atoms, max_mols=8192,
                   hbm_budget_bytes=args.serve_hbm_gib * 2 ** 30)
    srv32.run(mk(), realtime=False)   # untimed first pass, as for the BF16 server
    torch.cuda.synchronize()
    d32b = srv32.run(mk(), realtime=False)
    batched32 = atom_steps(d32b) / srv32.wall
    log(f"batched fp32 saturated: {srv32.wall:.2f} s, {batched32:.4g} atom-steps/s")
    del srv32
    # ... and one molecule per launch
    one = Server(eng32, max_atoms=max(len(z) for z, *_ in stream), max_mols=1)
    one.run(mk(4), realtime=False)
    torch.cuda.synchronize()
    d32 = one.run(mk(k), realtime=False)
    base_rate = atom_steps(d32) / one.wall
    log(f"launch-per-molecule fp32: {k} requests, {atom_steps(d32)} atom-steps in {one.wall:.2f} s, "
        f"{base_rate:.4g} atom-steps/s")
    return batched32, base_rate


def run_serving(args, stream):
    """Serve the request stream: realtime arrivals (latency), all requests at
    t = 0 (saturated throughput), and the plain-FP32 path one molecule per
    launch on the first requests of the stream (baseline)."""
    import torch
    from oracle import oracle as O
    from tensormol_amd.engine import BPEEEngine
    from tensormol_amd.model import BPEEModel
    from tensormol_amd.serving import Request, Server, atom_steps, latency_stats

    def mk(k=None):
        sel = stream if k is None else stream[:k]
        return [Request(r, Z, x, kind, st, t, seed=sd) for r, (Z, x, kind, st, t, sd) in enumerate(sel)]

    log = lambda m: print(f"[serve] {m}", file=sys.stderr, flush=True)
    model = BPEEModel([1, 6, 7, 8], dsf_alpha=0.2, seed=0)
    # angular rows of 128 entries (the loops run over the actual counts, so a
    # normal molecule pays nothing for them); requests that still overflow are
    # answered by the server's wide-row (256 / 128) engine
    eng = BPEEEngine(model, precision=args.precision, lda=128, check_overflow=False)
    srv = Server(eng, max_atoms=args.serve_atoms, max_mols=8192,
                 hbm_budget_bytes=args.serve_hbm_gib * 2 ** 30,
                 repack_every=args.serve_repack_every)
    srv.run(mk(64), realtime=False)   # warm-up
    torch.cuda.synchronize()
    done = srv.run(mk(), realtime=True)
    lat = latency_stats(done)
    rt_rate = atom_steps(done) / srv.wall
    rt_iters, rt_wall = srv.iters, srv.wall
    rt_rescued = srv.rescued   # row overflows answered by the wide-row server
    log(f"realtime: {len(done)} requests in {rt_wall:.2f} s, {rt_iters} iterations, "
        f"p50 {lat['p50_ms']:.2f} ms p99 {lat['p99_ms']:.2f} ms, {rt_rate:.4g} atom-steps/s")
    # saturated: the first pass at full batch size runs ~20 % slower (warm-up of
    # the full-size buffers and allocations); the second is timed
    srv.run(mk(), realtime=False)
    torch.cuda.synchronize()
    done_sat = srv.run(mk(), realtime=False)
    sat_rate = atom_steps(done_sat) / srv.wall
    log(f"saturated: {srv.wall:.2f} s, {srv.iters} iterations, {sat_rate:.4g} atom-steps/s")
    # force error of the single points against the FP64 oracle
    sp = sorted((r for r in done if r.kind == "sp"), key=lambda r: r.rid)[:args.serve_check]
    errs = []
    for r in sp:
        ref = O.evaluate(r.xyz, r.Z, np.array([0, len(r.Z)]), O.SymParams(model.eles),
                         model.energy_nets, model.charge_nets, alpha=model.alpha,
                         dsf_alpha=model.dsf_alpha, has_vdw=model.has_vdw)
        errs.append(float(np.sqrt(((r.forces / 2625499.638 + ref["gradient"]) ** 2).mean())))
    k = args.serve_baseline_requests
    batched32 = base_rate = None
    if args.serve_no_baselines:
        k = 0
    else:
        batched32, base_rate = serving_fp32_baselines(args, model, mk, stream, k, log)
    natoms = np.array([len(z) for z, *_ in stream])
    failed = {}
    for r in done:
        if r.error is not None:
            failed[r.kind] = failed.get(r.kind, 0) + 1
    return {
        "workload": f"{len(stream)} C/H/N/O chain requests, {natoms.min()}-{natoms.max()} atoms "
                    f"(mean {natoms.mean():.1f}), 50% single point / 30% 50-step NVT / 20% "
                    f"100-step FIRE, Poisson arrivals at {args.serve_rate:g}/s",
        "p50_ms": round(lat["p50_ms"], 3), "p99_ms": round(lat["p99_ms"], 3),
        "mean_ms": round(lat["mean_ms"], 3),
        "stream_atom_steps_per_s": round(rt_rate, 1), "stream_wall_s": round(rt_wall, 3),
        "stream_iterations": rt_iters,
        "saturated_atom_steps_per_s": round(sat_rate, 1),
        "launch_per_molecule_fp32_atom_steps_per_s": None if base_rate is None else round(base_rate, 1),
        "batched_fp32_saturated_atom_steps_per_s": None if batched32 is None else round(batched32, 1),
        "baseline_sample": (f"first {k} requests of the stream, served one molecule per launch "
                            "by the plain FP32 path") if base_rate is not None else None,
        "speedup_vs_launch_per_molecule_fp32": None if base_rate is None else round(sat_rate / base_rate, 2),
        "repack_every": args.serve_repack_every,
        "capacity_atoms": srv.max_atoms,
        "failed_requests_by_kind": failed,
        "rescued_requests": rt_rescued,
        "force_rms_err_vs_fp64_max": max(errs) if errs else None,
        "force_check_sample": f"first {len(errs)} single-point requests",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="bf16x3f")
    ap.add_argument("--nside", type=int, default=16)
    ap.add_argument("--edge", type=float, default=49.7)
    ap.add_argument("--dt", type=float, default=0.5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--steps-per-graph", type=int, default=10,
                    help="MD steps captured into one HIP graph (a divisor of --steps)")
    ap.add_argument("--serve-requests", type=int, default=4096, help="0 disables the serving leg")
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-atoms", type=int, default=65536)
    ap.add_argument("--serve-hbm-gib", type=float, default=16.0)
    ap.add_argument("--serve-baseline-requests", type=int, default=128)
    ap.add_argument("--serve-check", type=int, default=8)
    ap.add_argument("--serve-no-baselines", action="store_true",
                    help="skip the plain FP32 serving baselines (e.g. to trace the server alone)")
    ap.add_argument("--serve-repack-every", type=int, default=4,
                    help="repack a changed batch at most every K iterations (ghosts between); "
                         "4 measured best for p99 at 2000 req/s")
    ap.add_argument("--serve-only", action="store_true", help="print the serving record only")
    ap.add_argument("--calib-copy-mib", type=int, default=256,
                    help="size of the calibration copy run before the MD (0: none)")
    ap.add_argument("--train", action="store_true",
                    help="force-matching training steps on 256 chain molecules instead (own JSON line)")
    ap.add_argument("--train-mols", type=int, default=256)
    ap.add_argument("--train-steps", type=int, default=20)
    ap.add_argument("--train-no-graph", action="store_true",
                    help="time eager training steps only (no HIP-graph capture)")
    ap.add_argument("--sphere", action="store_true",
                    help="single-point E+F of the 24,000-atom C/H/N/O sphere instead (own JSON line)")
    ap.add_argument("--sphere-atoms", type=int, default=24000)
    ap.add_argument("--sphere-precision", default="fp16")
    ap.add_argument("--sphere-evals", type=int, default=50)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path (asked for "
              f"{args.gpus}); the single-GPU value is the figure of merit", file=sys.stderr)
        sys.exit(2)
    if args.sphere:
        print(json.dumps(run_sphere(args)), flush=True)
        return
    if args.train:
        print(json.dumps(run_train(args)), flush=True)
        return
    stream = None
    if args.serve_requests > 0:
        # generated by spawned worker processes before this process touches the GPU
        from tensormol_amd.workloads import request_stream
        t0 = time.perf_counter()
        stream = request_stream(args.serve_requests, rate=args.serve_rate)
        print(f"[serve] stream of {len(stream)} requests generated in {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
    import torch
    from tensormol_amd.engine import BPEEEngine
    if args.serve_only:
        print(json.dumps(run_serving(args, stream)), flush=True)
        return

    Z, x, box, model = build_system(args)
    N = len(Z)
    if args.calib_copy_mib > 0:
        # a plain device copy of known size: the calibration kernel for the
        # HBM byte counters (FETCH_SIZE / WRITE_SIZE) of profiled runs
        nb = args.calib_copy_mib << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    eng = BPEEEngine(model, precision=args.precision, check_overflow=False)
    ws = eng.prepare(Z, np.array([0, N]), box[None])
    print(f"[bench] {N} atoms, precision {args.precision}, warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)
    res = run_md(eng, ws, Z, x, box, args.steps, args.warmup, args.dt, graph=not args.no_graph,
                 steps_per_graph=args.steps_per_graph)
    value = N * args.steps / res["wall"]
    ms_per_step = 1e3 * res["wall"] / args.steps
    print(f"[bench] {value:.4g} atom-steps/s, {ms_per_step:.3f} ms/step", file=sys.stderr, flush=True)

    # frame-0 force error of this precision vs the FP64 oracle (+ CPU baseline timing)
    pos0 = torch.tensor(x, dtype=torch.float64, device=eng.device)
    eng.run(ws, pos0, check=True)
    torch.cuda.synchronize()
    r_gpu = eng.results(ws)
    side = torch.cuda.Stream(device=eng.device)
    side.wait_stream(torch.cuda.current_stream(eng.device))
    stages = stage_times(eng, ws, pos0, 10, side)
    P_r, P_r3 = int(ws.ncnt[:N].sum()), int(ws.acnt[:N].sum())
    print(f"[bench] stages (ms, eager) {stages}", file=sys.stderr, flush=True)
    cpu = None
    force_rms_err = None
    energy_err = None
    if not args.no_cpu_baseline:
        cores = os.cpu_count() or 1
        ref, cpu_s = cpu_baseline(Z, x, box, model, threads=cores)
        force_rms_err = float(np.sqrt(((r_gpu["gradient"] - ref["gradient"]) ** 2).mean()))
        energy_err = float(abs(r_gpu["Etotal"][0] - ref["Etotal"][0]))
        _, cpu1_s = cpu_baseline(Z, x, box, model, threads=1)
        cpu = {"value": N / cpu_s, "unit": "atom-steps/s", "cores": cores,
               "kind": "port",
               "sample": f"one FP64 E+F of the same {N}-atom box (neighbour lists included) by the "
                         f"numpy/torch-CPU oracle (a restatement of the reference's formulas, not "
                         f"the reference), {cpu_s:.1f} s on {cores} threads",
               "single_thread": {"value": N / cpu1_s, "unit": "atom-steps/s", "cores": 1,
                                 "sample": f"the same E+F on one thread, {cpu1_s:.1f} s"},
               "cpu_model": cpu_model_name()}
        print(f"[bench] oracle {cpu_s:.1f} s ({cores} threads), {cpu1_s:.1f} s (1 thread); force "
              f"rms err {force_rms_err:.3e} Ha/A; |dE| {energy_err:.3e} Ha", file=sys.stderr,
              flush=True)

    # plain FP32 path (VALU GEMMs, every intermediate in HBM) on the same box
    fp32_value = None
    if not args.no_fp32 and args.precision != "fp32":
        eng32 = BPEEEngine(model, precision="fp32", lda=128, check_overflow=False)
        ws32 = eng32.prepare(Z, np.array([0, N]), box[None])
        # same warmup, step count and HIP-graph replay as the timed run: the
        # random-weight system heats, so a shorter or earlier segment is cheaper
        r32 = run_md(eng32, ws32, Z, x, box, args.steps, args.warmup, args.dt,
                     graph=not args.no_graph, mlp_samples=0, steps_per_graph=args.steps_per_graph)
        fp32_value = N * args.steps / r32["wall"]
        print(f"[bench] plain fp32 path {fp32_value:.4g} atom-steps/s", file=sys.stderr, flush=True)

    serving = run_serving(args, stream) if stream is not None else None

    mlp_avg_ms = float(res["mlp_ms"].mean())
    flops = mlp_flops_per_atom(eng.D, int(model.hidden[0])) * N
    achieved = flops / (mlp_avg_ms * 1e-3) / 1e12
    peak = MLP_PEAK_TFLOPS[args.precision]
    # the MLP stage is one launch per tile size in use (64-row tiles in whole
    # rounds over the CUs, then 32-row tiles for the rest); its time is the
    # span of those launches on their stream
    mlp_launches = [f"mlp_mfma_kernel<{eng.mfma.mode}, {r}>" for r, n in
                    ((64, ws.mlp_ws["n64"]), (32, ws.mlp_ws["n32"])) if n > 0] \
        if args.precision != "fp32" else ["mlp_f32 kernels"]
    mlp_roof = {"bound": "mfma" if args.precision != "fp32" else "valu", "kernel": "mlp_mfma_kernel",
                "launches": mlp_launches,
                "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": None,
                "flops_per_launch": flops, "avg_launch_ms": round(mlp_avg_ms, 4)}
    # The Coulomb / vdW sweep over the unordered pairs within 15 A, VALU-bound,
    # ~50 FP32 operations per unordered pair (forward, potential and force),
    # against the FP32 vector peak.
    P_c = coulomb_pairs(x, box)
    coul_avg_ms = float(res["coul_ms"].mean())
    coul_flops = COUL_FLOPS_PER_PAIR * P_c
    coul_achieved = coul_flops / (coul_avg_ms * 1e-3) / 1e12
    coul_kernel = "coulomb_tab_kernel" if eng.coulomb_tables else "coulomb_kernel"
    coul_roof = {"bound": "valu", "kernel": coul_kernel, "launches": [coul_kernel],
                 "achieved": round(coul_achieved, 3),
                 "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": round(coul_achieved / FP32_VALU_PEAK_TFLOPS, 4), "traffic": None,
                 "flops_per_launch": coul_flops, "pairs_15A": P_c,
                 "avg_launch_ms": round(coul_avg_ms, 4)}
    # HBM bytes per launch come from two rocprofv3 counter runs (FETCH_SIZE,
    # WRITE_SIZE) calibrated on a plain copy; they cannot be counted live here
    tpath = os.path.join(ROOT, "profiles", f"water_box_12288_{args.precision}_hbm_traffic.json")
    if os.path.exists(tpath) and N == 12288:
        tr = json.load(open(tpath))
        kern = {k.split("<")[0]: v for k, v in tr["kernels"].items()}
        if coul_kernel in kern:
            coul_roof["traffic"] = kern[coul_kernel]["bytes"]
            coul_roof["traffic_source"] = tr["source"]
        if tr.get("mlp_precision") == args.precision:
            mlp_roof["traffic"] = tr["mlp_stage_bytes"]
            mlp_roof["traffic_source"] = tr["source"]
    # the headline roofline is the stage that takes the longest per step
    roofline = dict(mlp_roof if mlp_avg_ms >= coul_avg_ms else coul_roof)
    bdesc = descriptor_bytes(N, eng.D, P_r, P_r3)
    by_stage = {
        "coulomb": coul_roof,
        "mlp": mlp_roof,
        "descriptors": {"bound": "hbm", "unit": "GB/s", "peak": 8000.0, "bytes_per_launch": bdesc,
                        "achieved": round(bdesc / (stages["descriptors"] * 1e-3) / 1e9, 1)},
        "adjoint": {"bound": "hbm", "unit": "GB/s", "peak": 8000.0, "bytes_per_launch": bdesc + 12 * N,
                    "achieved": round((bdesc + 12 * N) / (stages["adjoint"] * 1e-3) / 1e9, 1)},
    }
    for v in by_stage.values():
        v["frac"] = round(v["achieved"] / v["peak"], 4)
    # total energy (Ha) at the start and end of the timed steps; KE sums are
    # m v^2 / 2 in kg/mol A^2/fs^2 = 1e10 J/mol
    etot = lambda e: e[0] + e[1] * 1e10 / 2625499.638
    drift = etot(res["e_end"]) - etot(res["e_start"])
    out = {
        "metric": "atom-steps/s (E+F) at force RMS<1e-3 Ha/A, BF16-MFMA vs plain FP32; p99 ms",
        "value": round(value, 1),
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "p50_ms": None if serving is None else serving["p50_ms"],
        "p99_ms": None if serving is None else serving["p99_ms"],
        "md_step_p50_ms": round(float(np.percentile(res["per_step_ms"], 50)), 4),
        "md_step_p99_ms": round(float(np.percentile(res["per_step_ms"], 99)), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": MFMA_DTYPE[args.precision],
        "data": "synthetic",
        "config": {"workload": f"periodic {N // 3}-water box NVE MD ({N} atoms, {args.edge} A cube), "
                               "BP+EE 256-512-512-512-1 x2 nets, DSF 15 A",
                   "atoms": N, "dt_fs": args.dt, "parallelism": "single GPU",
                   "steps_per_graph": res["steps_per_graph"],
                   "hip_graph": not args.no_graph,
                   "mlp_precision": args.precision,
                   "mlp_precision_note": MLP_PRECISION_NOTE.get(args.precision, "")},
        "nve_energy_drift_Ha": round(float(drift), 6),
        "force_rms_err_vs_fp64": force_rms_err,
        "energy_err_vs_fp64_Ha": energy_err,
        "plain_fp32_value": None if fp32_value is None else round(fp32_value, 1),
        "speedup_vs_plain_fp32": None if fp32_value is None else round(value / fp32_value, 2),
        "serving": serving,
        "stage_ms_eager": {k: round(v, 4) for k, v in stages.items()},
        "pairs_4.6A": P_r, "pairs_3.1A": P_r3,
        "roofline_by_stage": by_stage,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
