"""Continuous-batching server: results vs the FP64 oracle, independence of a
request's result from the batch it is served in, and admission under a small
capacity.

Single points are checked against the oracle at the engine's end-to-end gates
in its default precision (bf16x3f): force RMS error < 1e-3 Ha/A and
|dE| < 1e-4 Ha per molecule.  Every per-molecule computation of the
engine is ordered by molecule-local indices only (candidate sweeps start at the
molecule's first atom, per-molecule reductions use a fixed tree over local
offsets, MFMA rows are independent), so a request served alone and inside a
mix must give bit-identical positions and forces.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tensormol_amd.params import JOULEPERHARTREE
from tests import helpers as H

pytestmark = pytest.mark.gpu

KINDS = ((0.4, "sp", 1), (0.3, "nvt", 6), (0.3, "opt", 12))


def _stream(n=24, seed=21):
    from tensormol_amd.serving import make_stream
    return make_stream(n, lo=8, hi=64, seed=seed, kinds=KINDS, workers=1)


def _copy(reqs):
    from tensormol_amd.serving import Request
    return [Request(r.rid, r.Z, r.xyz, r.kind, r.nsteps, r.t_arrival, seed=r.seed) for r in reqs]


def _oracle(model, Z, x):
    sp = O.SymParams(model.eles)
    return O.evaluate(x, Z, np.array([0, len(Z)]), sp, model.energy_nets, model.charge_nets,
                      alpha=model.alpha, dsf_alpha=model.dsf_alpha, has_vdw=model.has_vdw,
                      need_forces=True)


@pytest.fixture(scope="module")
def served(cuda):
    from tensormol_amd.engine import BPEEEngine
    from tensormol_amd.serving import Server
    model = H.chno_model()
    eng = BPEEEngine(model, precision="bf16x3f")
    srv = Server(eng, max_atoms=4096, max_mols=256)
    reqs = _stream()
    done = srv.run(reqs, realtime=False)
    torch.cuda.synchronize()
    return model, eng, srv, reqs, {r.rid: r for r in done}


def test_all_requests_served(served):
    model, eng, srv, reqs, done = served
    assert sorted(done) == [r.rid for r in reqs]
    for r in reqs:
        d = done[r.rid]
        assert d.error is None
        if r.kind == "sp":
            assert d.steps_done == 1
        if r.kind == "nvt":
            assert d.steps_done == r.nsteps
        if r.kind == "opt":
            assert 1 <= d.steps_done <= r.nsteps
    print(f"iterations {srv.iters}, atom-steps {srv.atom_steps}")


def test_single_points_vs_oracle(served):
    model, eng, srv, reqs, done = served
    n = 0
    for r in reqs:
        if r.kind != "sp":
            continue
        d = done[r.rid]
        ref = _oracle(model, r.Z, r.xyz)
        f_ref = -ref["gradient"]
        frms = np.sqrt(((d.forces / JOULEPERHARTREE - f_ref) ** 2).mean())
        dE = abs(d.energy - ref["Etotal"][0])
        print(f"sp rid {r.rid} natom {len(r.Z)}: |dE| {dE:.2e} Ha, force rms err {frms:.2e} Ha/A")
        assert frms < 1e-3
        assert dE < 1e-4
        n += 1
    assert n > 0


def test_fire_lowers_energy(served):
    model, eng, srv, reqs, done = served
    n = 0
    for r in reqs:
        if r.kind != "opt":
            continue
        d = done[r.rid]
        e0 = _oracle(model, r.Z, r.xyz)["Etotal"][0]
        e1 = _oracle(model, r.Z, d.positions)["Etotal"][0]
        print(f"opt rid {r.rid} natom {len(r.Z)} steps {d.steps_done}: E {e0:.6f} -> {e1:.6f} Ha")
        assert e1 < e0
        n += 1
    assert n > 0


def test_alone_equals_mix(served):
    """A request's result does not depend on the other requests in its batch."""
    from tensormol_amd.serving import Server
    model, eng, srv, reqs, done = served
    picks = {}
    for r in reqs:
        picks.setdefault(r.kind, [])
        if len(picks[r.kind]) < 2:
            picks[r.kind].append(r)
    alone_srv = Server(eng, max_atoms=256, max_mols=4)
    for kind, rs in picks.items():
        for r in rs:
            (a,) = alone_srv.run(_copy([r]), realtime=False)
            d = done[r.rid]
            assert a.steps_done == d.steps_done
            if kind == "sp":
                np.testing.assert_array_equal(a.forces, d.forces)
                assert a.energy == d.energy
            else:
                np.testing.assert_array_equal(a.positions, d.positions)


def test_admission_under_small_capacity(served):
    """With room for a few molecules at a time the server queues requests and
    gives the same results."""
    from tensormol_amd.serving import Server
    model, eng, srv, reqs, done = served
    small = Server(eng, max_atoms=160, max_mols=16)
    out = {r.rid: r for r in small.run(_copy(reqs), realtime=False)}
    assert sorted(out) == sorted(done)
    assert small.iters > srv.iters
    for rid, d in done.items():
        o = out[rid]
        if d.kind == "sp":
            np.testing.assert_array_equal(o.forces, d.forces)
        else:
            np.testing.assert_array_equal(o.positions, d.positions)


def test_capacity_guard(served):
    from tensormol_amd.serving import Server
    model, eng, srv, reqs, done = served
    big = max(reqs, key=lambda r: len(r.Z))
    tiny = Server(eng, max_atoms=len(big.Z) - 1, max_mols=4)
    with pytest.raises(ValueError):
        tiny.run(_copy([big]), realtime=False)


def _dense_molecule(n=60, radius=2.0, seed=5):
    """60 C/H atoms in a 2 A ball (min distance 0.5 A): 59 neighbours within
    4.6 A (inside the 128-entry rows) and up to 59 within 3.1 A (beyond the
    48-entry angular rows), as a collapsed geometry would have."""
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.uniform(-radius, radius, 3)
        if np.linalg.norm(p) <= radius and all(np.linalg.norm(p - q) > 0.5 for q in pts):
            pts.append(p)
    Z = np.where(np.arange(n) % 2 == 0, 6, 1).astype(np.int32)
    return Z, np.array(pts)


def test_failure_stays_with_its_molecule(served):
    """A molecule whose rows overflow fails alone; its batch-mates' results are
    the same as in any other batch."""
    from tensormol_amd.serving import Request, Server
    model, eng, srv, reqs, done = served
    normal = [r for r in reqs if r.kind == "sp"][:2] + [r for r in reqs if r.kind == "nvt"][:2]
    Zd, xd = _dense_molecule()
    bad = Request(10 ** 6, Zd, xd, "nvt", 6, 0.0, seed=1)
    s2 = Server(eng, max_atoms=1024, max_mols=8, rescue_rows=None)
    out = {r.rid: r for r in s2.run(_copy(normal) + [bad], realtime=False)}
    assert out[bad.rid].error is not None and out[bad.rid].steps_done == 1
    assert out[bad.rid].positions is None
    for r in normal:
        o, d = out[r.rid], done[r.rid]
        assert o.error is None
        if r.kind == "sp":
            np.testing.assert_array_equal(o.forces, d.forces)
        else:
            np.testing.assert_array_equal(o.positions, d.positions)


def test_row_overflow_rescued_by_wide_rows(served):
    """A molecule whose angular rows overflow the 48-entry rows is answered by
    the wide-row (256 / 128) server: the same result as that server gives it
    alone, bit for bit, while its batch-mates keep their results."""
    from tensormol_amd.engine import BPEEEngine
    from tensormol_amd.serving import Request, Server
    model, eng, srv, reqs, done = served
    normal = [r for r in reqs if r.kind == "sp"][:2] + [r for r in reqs if r.kind == "opt"][:2]
    Zd, xd = _dense_molecule()
    dense = [Request(10 ** 6, Zd, xd, "nvt", 6, 0.0, seed=1),
             Request(10 ** 6 + 1, Zd, xd, "sp", 1, 0.0)]
    s2 = Server(eng, max_atoms=1024, max_mols=8)
    out = {r.rid: r for r in s2.run(_copy(normal) + _copy(dense), realtime=False)}
    torch.cuda.synchronize()
    assert sorted(out) == sorted([r.rid for r in normal] + [r.rid for r in dense])
    for r in normal:
        o, d = out[r.rid], done[r.rid]
        assert o.error is None and not o.rescued
        if r.kind == "sp":
            np.testing.assert_array_equal(o.forces, d.forces)
        else:
            np.testing.assert_array_equal(o.positions, d.positions)
    wide = Server(BPEEEngine(model, precision=eng.precision, ldr=256, lda=128), max_atoms=1024,
                  max_mols=8, rescue_rows=None)
    alone = {r.rid: r for r in wide.run(_copy(dense), realtime=False)}
    torch.cuda.synchronize()
    for r in dense:
        o, a = out[r.rid], alone[r.rid]
        assert o.rescued and o.error is None and a.error is None, (o.error, a.error)
        assert o.steps_done == r.nsteps and o.energy == a.energy
        if r.kind == "sp":
            np.testing.assert_array_equal(o.forces, a.forces)
        else:
            np.testing.assert_array_equal(o.positions, a.positions)
    assert o.t_done is not None


def test_rescued_single_point_vs_oracle(cuda):
    """The collapsed molecule's energy and forces from the wide-row server (FP32
    path) against the FP64 oracle."""
    from tensormol_amd.engine import BPEEEngine
    from tensormol_amd.serving import Request, Server
    model = H.chno_model()
    Zd, xd = _dense_molecule()
    srv = Server(BPEEEngine(model, precision="fp32"), max_atoms=256, max_mols=4)
    (r,) = srv.run([Request(7, Zd, xd, "sp", 1, 0.0)], realtime=False)
    torch.cuda.synchronize()
    assert r.rescued and r.error is None
    ref = _oracle(model, Zd, xd)
    dE = abs(r.energy - ref["Etotal"][0])
    F_ref = -ref["gradient"] * JOULEPERHARTREE
    frms = np.sqrt(((r.forces - F_ref) ** 2).mean()) / JOULEPERHARTREE
    fscale = np.sqrt((F_ref ** 2).mean()) / JOULEPERHARTREE
    print(f"dense 60-atom ball: |dE| {dE:.2e} Ha ({dE / len(Zd):.2e} per atom), force rms err "
          f"{frms:.2e} Ha/A (rms {fscale:.2e})")
    assert dE / len(Zd) < 2e-6
    assert frms < 1e-4 * max(1.0, fscale)


def test_lazy_repack_same_results(served):
    """With repack_every = 4 (departed requests ride along as discarded ghosts,
    arrivals wait for the next repack) every request ends with the same steps,
    energy and coordinates/forces, bit for bit."""
    from tensormol_amd.serving import Server
    model, eng, srv, reqs, done = served
    lazy = Server(eng, max_atoms=4096, max_mols=256, repack_every=4)
    got = {r.rid: r for r in lazy.run(_copy(reqs), realtime=False)}
    torch.cuda.synchronize()
    assert sorted(got) == sorted(done)
    assert lazy.iters > 0
    for rid, d in done.items():
        g = got[rid]
        assert g.error == d.error and g.steps_done == d.steps_done, rid
        assert g.energy == d.energy, rid
        if d.kind == "sp":
            assert np.array_equal(g.forces, d.forces), rid
        else:
            assert np.array_equal(g.positions, d.positions), rid


def test_graph_replayed_steps_same_results(served):
    """Relaunched layouts replayed as HIP graphs (captured on a layout's first
    relaunch) give every request the same steps, energy and coordinates /
    forces, bit for bit, as eager launches; the eager program work between
    replays (repacks, result copies) runs as usual."""
    from tensormol_amd.serving import Server
    model, eng, srv, reqs, done = served
    eager = Server(eng, max_atoms=4096, max_mols=256, repack_every=4)
    ref = {r.rid: r for r in eager.run(_copy(reqs), realtime=False)}
    g = Server(eng, max_atoms=4096, max_mols=256, repack_every=4, graphs=True)
    got = {r.rid: r for r in g.run(_copy(reqs), realtime=False)}
    torch.cuda.synchronize()
    print(f"graph captures {g.graph_captures}, replays {g.graph_replays}, iterations {g.iters}")
    assert g.graph_replays > g.graph_captures > 0
    assert sorted(got) == sorted(ref)
    for rid, d in ref.items():
        x = got[rid]
        assert x.error == d.error and x.steps_done == d.steps_done, rid
        assert x.energy == d.energy, rid
        if d.kind == "sp":
            assert np.array_equal(x.forces, d.forces), rid
        else:
            assert np.array_equal(x.positions, d.positions), rid


def test_packed_outputs_copy_every_part(cuda):
    """tm_pack_copy through PackedOutputs: each part lands at its own offset,
    unchanged, including more parts than one launch takes and empty parts."""
    from tensormol_amd.staging import PackedOutputs
    g = torch.Generator(device="cpu").manual_seed(5)
    srcs = {}
    for i in range(21):
        n = int(torch.randint(0, 700, (1,), generator=g))
        if i % 3 == 0:
            srcs[f"i{i}"] = torch.randint(-2 ** 31, 2 ** 31 - 1, (n,), generator=g, dtype=torch.int32)
        else:
            srcs[f"d{i}"] = torch.randn(n, 3, generator=g, dtype=torch.float64)
    srcs["empty"] = torch.zeros(0, 3, dtype=torch.float64)
    total = sum(t.numel() * t.element_size() + 8 for t in srcs.values())
    outs = PackedOutputs(cuda, total, max_parts=32)
    for rep in range(2):     # a second, smaller set reuses the buffers
        dev = {k: v.to(cuda) for k, v in srcs.items()}
        outs.begin()
        for k, v in dev.items():
            outs.add(k, v)
        outs.mark()
        outs.wait()
        for k, v in srcs.items():
            np.testing.assert_array_equal(outs.host(k), v.numpy(), err_msg=k)
        srcs = {k: v[: len(v) // 2] for k, v in list(srcs.items())[:7]}
