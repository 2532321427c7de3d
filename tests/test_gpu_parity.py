"""GPU parity of the HIP path against the FP64 oracle (through the C ABI).

Bars: neighbour pair and triple sets bit-exact (integer data); descriptors
within 2e-5 relative (FP32 arithmetic); energies, charges and forces within
the tolerances stated per test.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

SYSTEMS = [("water64", "water"), ("chains", "chno"), ("waterbox10", "water"), ("chno_small", "chno"),
           ("morphine", "chno"), ("water10", "water"), ("peptide188", "chno"), ("johnsonmols", "chno")]


def _engine(modelname, precision="fp32", **kw):
    from tensormol_amd.engine import BPEEEngine
    return BPEEEngine(H.model_for(modelname), precision=precision, **kw)


def _prepare(eng, name, dev):
    Z, x, ms, boxes = H.system(name)
    ws = eng.prepare(Z, ms, boxes, pos0=x)
    pos = torch.tensor(x, dtype=torch.float64, device=dev)
    return ws, pos


def gpu_pairs_triples(eng, ws):
    """Rows of the GPU lists as reference-style arrays with local indices."""
    ldr, lda, nele = eng.ldr, eng.lda, eng.nele
    nbr = ws.nbr.cpu().numpy().reshape(ws.N, ldr)
    ncnt = ws.ncnt.cpu().numpy()
    acnt = ws.acnt.cpu().numpy()
    ms = ws.ms_np
    mol = ws.mol_np
    elem = ws.elem_np
    rows = []
    for i in range(ws.N):
        m = mol[i]
        for n in range(ncnt[i]):
            j = nbr[i, n]
            rows.append((m, i - ms[m], j - ms[m], elem[j]))
    trp = np.array(rows, dtype=np.int64).reshape(-1, 4)
    ntrip = acnt.astype(np.int64) * (acnt.astype(np.int64) - 1) // 2
    start = np.concatenate([[0], np.cumsum(ntrip)]).astype(np.int64)
    out = torch.zeros(max(1, int(start[-1])) * 4, dtype=torch.int32, device=ws.nbr.device)
    ts = torch.tensor(start[:-1], dtype=torch.int64, device=ws.nbr.device)
    from tensormol_amd._lib import ptr
    eng.lib.tm_triples(ws.N, ldr, lda, nele, ptr(ws.nbr), ptr(ws.ang), ptr(ws.acnt), ptr(ws.aseg),
                       ptr(ts), ptr(out), eng._stream())
    o = out.cpu().numpy().reshape(-1, 4)[: start[-1]]
    ci = np.repeat(np.arange(ws.N), ntrip)
    cm = mol[ci]
    trt = np.stack([cm, ci - ms[cm], o[:, 0] - ms[cm], o[:, 1] - ms[cm], o[:, 2]], axis=1)
    mil = np.stack([cm, ci - ms[cm], o[:, 2], o[:, 3]], axis=1)
    return trp, trt, mil


def test_flags_kept_without_clear(cuda):
    """build_lists(clear_flags=False) keeps the status flags of earlier builds
    (steps captured into one graph accumulate them; the check after a replay
    sees every step's), and the default clears them.  The overflow comes from
    the same periodic box with its coordinates halved: eight times the density,
    rows longer than their 128 slots (a flagged, clipped case)."""
    eng = _engine("water")
    ws, pos = _prepare(eng, "waterbox10", cuda)
    Z, x, ms, boxes = H.system("waterbox10")
    lo = x.min(axis=0)
    dense = torch.tensor(lo + 0.5 * (x - lo), dtype=torch.float64, device=cuda)
    eng.build_lists(ws, dense)
    f1 = ws.flags.cpu().numpy().copy()
    assert f1[0] != 0
    eng.build_lists(ws, pos, clear_flags=False)
    np.testing.assert_array_equal(ws.flags.cpu().numpy(), f1)
    eng.build_lists(ws, pos)
    assert not ws.flags.cpu().numpy().any()
    eng.check_flags(ws)


@pytest.mark.parametrize("name,modelname", SYSTEMS)
def test_neighbour_sets_bit_exact(cuda, name, modelname):
    eng = _engine(modelname)
    ws, pos = _prepare(eng, name, cuda)
    eng.build_lists(ws, pos)
    eng.check_flags(ws)
    torch.cuda.synchronize()
    trp, trt, mil = gpu_pairs_triples(eng, ws)
    Z, x, ms, boxes = H.system(name)
    B = len(ms) - 1
    maxn = int(np.diff(ms).max())
    xyzs = np.zeros((B, maxn, 3)); Zs = np.zeros((B, maxn), dtype=np.int64)
    for m in range(B):
        n = ms[m + 1] - ms[m]
        xyzs[m, :n] = x[ms[m]:ms[m + 1]]; Zs[m, :n] = Z[ms[m]:ms[m + 1]]
    rtrp, rtrt, rmil, _ = O.pairs_triples_with_ele(xyzs, np.diff(ms), Zs, eng.model.eles, 4.6, 3.1,
                                                    boxes=boxes)
    # GPU rows are emitted per centre in (l, j) order -> reference (m, i, l, j) order
    assert trp.shape == rtrp.shape, (trp.shape, rtrp.shape)
    np.testing.assert_array_equal(trp, rtrp)
    # triples: GPU emits per centre grouped by (p, k, j) -> same as the lexsort
    assert trt.shape == rtrt.shape, (trt.shape, rtrt.shape)
    np.testing.assert_array_equal(trt, rtrt)
    np.testing.assert_array_equal(mil, rmil.astype(np.int64))


@pytest.mark.parametrize("name,modelname", SYSTEMS)
def test_mirror_slots(cuda, name, modelname):
    """rev from the descriptor forward equals the neighbour build's own pass,
    and every slot (i -> j) points at the slot (j -> i)."""
    from tensormol_amd._lib import ptr
    eng = _engine(modelname)
    ws, pos = _prepare(eng, name, cuda)
    eng.build_lists(ws, pos)
    eng.descriptors(ws, pos)          # fills ws.rev
    eng.check_flags(ws)
    torch.cuda.synchronize()
    rev_fwd = ws.rev.clone()
    ws.rev.fill_(-7)
    L, st = eng.lib, eng._stream()
    g = eng._grid_args(ws)
    L.tm_nl_build(ws.N, ptr(pos), ptr(ws.elem), ptr(ws.mol), ptr(ws.mol_start),
                  ptr(ws.box), *g, ws.periodic, eng.nele, eng.model.sym.rr, eng.s_r, eng.s_a,
                  eng.ldr, eng.lda, ptr(ws.nbr), ptr(ws.ncnt), ptr(ws.rseg), ptr(ws.ang),
                  ptr(ws.acnt), ptr(ws.aseg), ptr(ws.rev), ptr(ws.flags), st)
    torch.cuda.synchronize()
    ncnt = ws.ncnt.cpu().numpy()
    valid = (np.arange(eng.ldr)[None, :] < ncnt[:, None]).reshape(-1)
    a = rev_fwd.cpu().numpy()[valid]
    b = ws.rev.cpu().numpy()[valid]
    np.testing.assert_array_equal(a, b)
    nbr = ws.nbr.cpu().numpy()
    centre = np.repeat(np.arange(ws.N), eng.ldr)[valid]
    np.testing.assert_array_equal(nbr[a], centre)


@pytest.mark.parametrize("name,modelname", SYSTEMS)
def test_descriptors(cuda, name, modelname):
    eng = _engine(modelname)
    ws, pos = _prepare(eng, name, cuda)
    eng.build_lists(ws, pos)
    eng.descriptors(ws, pos)
    G = ws.G.cpu().numpy().astype(np.float64)
    ref = H.oracle_eval(name, modelname, need_forces=False)["G"]
    err = np.abs(G - ref).max() / max(1.0, np.abs(ref).max())
    print(f"{name}: descriptor max rel err {err:.3e}, |G|max {np.abs(ref).max():.3f}")
    assert err < 2e-5


@pytest.mark.parametrize("name,modelname", SYSTEMS)
def test_energy_forces_fp32(cuda, name, modelname):
    eng = _engine(modelname, "fp32")
    ws, pos = _prepare(eng, name, cuda)
    eng.run(ws, pos)
    torch.cuda.synchronize()
    r = eng.results(ws)
    ref = H.oracle_eval(name, modelname)
    dE = np.abs(r["Etotal"] - ref["Etotal"])
    natom = np.diff(H.system(name)[2])
    frms = np.sqrt(((r["gradient"] - ref["gradient"]) ** 2).mean())
    fscale = np.sqrt((ref["gradient"] ** 2).mean())
    qerr = np.abs(r["charge"] - ref["charge"]).max()
    print(f"{name}: |dE| max {dE.max():.3e} Ha (per atom {np.max(dE / natom):.3e}), "
          f"force rms err {frms:.3e} Ha/A (rms {fscale:.3e}), q err {qerr:.3e}, "
          f"dEcc {np.abs(r['Ecc'] - ref['Ecc']).max():.3e} dEvdw {np.abs(r['Evdw'] - ref['Evdw']).max():.3e}")
    assert np.all(dE / natom < 2e-6)
    assert frms < 1e-4
    assert qerr < 1e-4
    np.testing.assert_allclose(r["dipole"], ref["dipole"], atol=5e-3 * np.sqrt(natom.max()))


@pytest.mark.parametrize("name,modelname", [("waterbox10", "water")])
def test_fused_abi_paths_bit_identical(cuda, name, modelname):
    """The engine's fused paths equal the separate ones bit for bit: charges
    written in cell order by tm_charges (tm_coulomb with xq = NULL) against the
    gather inside tm_coulomb, the molecule means formed inside tm_charges
    (mol_start given; they also land in qmean) against qmean as input, and the
    charge upstream formed inside tm_sym_adjoint (u = NULL, phi and phimean
    given) against tm_charge_upstream.  The in-launch means agree with
    tm_mol_sum_f32 to rounding."""
    from tensormol_amd._lib import P, ptr
    from tensormol_amd.params import PARAMS
    eng = _engine(modelname, "fp32", coulomb_tables=False)   # the analytic tm_coulomb sweep
    ws, pos = _prepare(eng, name, cuda)
    assert ws.use_cells
    eng.run(ws, pos)
    torch.cuda.synchronize()
    fused = {k: getattr(ws, k).clone() for k in ("phi", "gradc", "eatom", "wslot", "q")}
    L, st = eng.lib, eng._stream()
    g = eng._grid_args(ws)
    mean_ref = torch.full_like(ws.qmean, float("nan"))
    L.tm_mol_sum_f32(ws.B, ptr(ws.mol_start), 1, ptr(ws.qraw), 1, ptr(mean_ref), 1, ptr(ws.red),
                     ws.N, st)
    torch.cuda.synchronize()
    scale = ws.qraw[:ws.N].abs().double().mean().item()
    assert (ws.qmean[:ws.B] - mean_ref[:ws.B]).abs().max().item() <= 1e-13 * scale
    L.tm_charges(ws.N, ptr(ws.mol), None, ptr(ws.qmean), ptr(ws.qraw), ptr(pos), ptr(ws.xq),
                 ptr(ws.q), ptr(ws.qx), None, None, None, None, st)
    ws.xqc.zero_()
    ws.elc.zero_()
    L.tm_coulomb(ws.N, ptr(ws.xq), ptr(ws.elem), ptr(ws.mol), ptr(ws.mol_start), ptr(ws.box),
                 g[0], g[1], g[2], g[3], g[5], g[6], g[7], g[8], g[9], g[10], g[11], g[12], g[13],
                 g[14], ws.periodic, eng.dsf.ctypes.data_as(P),
                 eng.model.ee_off, eng.vdw_width, int(bool(PARAMS["AddEcc"])),
                 int(eng.model.has_vdw), eng.nele, eng.c6.ctypes.data_as(P),
                 eng.r0.ctypes.data_as(P), ptr(ws.xqc), ptr(ws.elc), ptr(ws.phi),
                 ptr(ws.gradc), ptr(ws.eatom), ptr(ws.flags), st)
    L.tm_charge_upstream(ws.N, ptr(ws.mol), ptr(ws.phi), ptr(ws.phimean), ptr(ws.u), st)
    sym = eng.model.sym
    from tensormol_amd._lib import fptr
    L.tm_sym_adjoint(ws.N, ptr(pos), ptr(ws.box), ptr(ws.mol), ptr(ws.elem), ptr(ws.nbr),
                     ptr(ws.ncnt), ptr(ws.rseg), ptr(ws.ang), ptr(ws.acnt), ptr(ws.aseg), eng.ldr,
                     eng.lda, eng.nele, sym.nr_r, sym.eta, sym.zeta, sym.rr, sym.ra,
                     fptr(eng.sym_rs_r), fptr(eng.sym_rs_a), fptr(eng.sym_th), ptr(ws.JE),
                     ptr(ws.JQ), ptr(ws.u), None, None, eng.D, ptr(ws.wslot), st)
    torch.cuda.synchronize()
    for k, v in fused.items():
        assert torch.equal(getattr(ws, k), v), k
