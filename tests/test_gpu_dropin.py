"""The drop-in Python API over the HIP library, against the FP64 oracle.

Covers the reference entry points a TensorMol script calls:
  - MolEmb.Make_NListNaive (C_API/MolEmb.cpp:1180-1247), NeighborListSet
    .buildPairsAndTriplesWithEleIndex / .buildPairs (Neighbors.py:262-282,
    344-423): integer arrays bit-exact (same float64 predicate), including
    pairs a few ulps either side of the cutoff, in open space and across a
    periodic wrap;
  - TFMolManage.EvalBPDirectEEUpdateSingle / Set / SinglePeriodic
    (TFMolManage.py:1260-1358) and the instance's positional evaluate
    (TFMolInstanceDirect.py:5684-5711): tuple shapes, units (forces
    -JOULEPERHARTREE * dE/dx in J/mol/A), values within the FP32 path's
    tolerance of the oracle, HasVdw semantics;
  - the host drivers VelocityVerlet.Prop and GeomOptimizer.Opt over a closure
    that calls the manager.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

FP32_E_TOL = 2e-6     # Ha per atom, the FP32 path against the oracle
FP32_G_TOL = 1e-4     # Ha/A, force RMS per molecule (FP32 descriptors and adjoint)


def _padded(name):
    Z, x, ms, boxes = H.system(name)
    B = len(ms) - 1
    natom = np.diff(ms)
    maxn = int(natom.max())
    xyzs = np.zeros((B, maxn, 3))
    Zs = np.zeros((B, maxn), dtype=np.int32)
    for m in range(B):
        xyzs[m, :natom[m]] = x[ms[m]:ms[m + 1]]
        Zs[m, :natom[m]] = Z[ms[m]:ms[m + 1]]
    return xyzs, Zs, natom, boxes


# ------------------------------------------------------------ neighbour API
@pytest.mark.parametrize("doperms", [1, 0])
def test_make_nlist_naive(cuda, doperms):
    from tensormol_amd.neighbors import Make_NListNaive
    Z, x = H.system("water64")[:2]
    nreal = 150                                   # atoms 150.. act as images
    got = Make_NListNaive(x, 4.6, nreal, doperms)
    ref = O.nlist_naive(x, 4.6, nreal, doperms)
    assert len(got) == nreal
    for i in range(nreal):
        assert got[i] == [int(j) for j in ref[i]], i


@pytest.mark.parametrize("name", ["chains", "water10", "johnsonmols"])
def test_neighborlistset_with_ele(cuda, name):
    from tensormol_amd.neighbors import NeighborListSet
    xyzs, Zs, natom, _ = _padded(name)
    model = H.model_for("water" if name == "water10" else "chno")
    eles = np.asarray(model.eles).reshape(-1, 1)
    elep = np.asarray(model.sym.epairs)
    NL = NeighborListSet(xyzs, natom, True, True, Zs, sort_=True)
    trpE, trtE, mil_jk, jk_max = NL.buildPairsAndTriplesWithEleIndex(4.6, 3.1, eles, elep)
    rp, rt, rm, rj = O.pairs_triples_with_ele(xyzs, natom, Zs, model.eles, 4.6, 3.1)
    assert trpE.dtype == np.uint64 and trtE.dtype == np.uint64 and mil_jk.dtype == np.float64
    np.testing.assert_array_equal(trpE.astype(np.int64), rp)
    np.testing.assert_array_equal(trtE.astype(np.int64), rt)
    np.testing.assert_array_equal(mil_jk, rm)
    assert jk_max == rj
    # the 15 A electrostatic half list, each pair once under the lower index
    NLEE = NeighborListSet(xyzs, natom, False, False, None)
    p = NLEE.buildPairs(15.0)
    np.testing.assert_array_equal(p.astype(np.int64), O.half_pairs(xyzs, natom, 15.0))
    # full list without element tags
    NLf = NeighborListSet(xyzs, natom, False, True, Zs)
    pf = NLf.buildPairs(4.6).astype(np.int64)
    np.testing.assert_array_equal(pf, rp[np.lexsort((rp[:, 2], rp[:, 1], rp[:, 0]))][:, :3])


def _ulp_distances(rc, k=6):
    d0 = rc - 1e-13
    out = [d0]
    lo = hi = d0
    for _ in range(k):
        lo = float(np.nextafter(lo, 0.0))
        hi = float(np.nextafter(hi, np.inf))
        out += [lo, hi]
    return sorted(out)


def test_cutoff_ulp_pairs_open(cuda):
    """Pairs whose separations step one ulp at a time across the cutoff: the
    GPU keeps exactly the pairs the float64 predicate keeps (Make_NListNaive
    and the 15 A half list)."""
    from tensormol_amd.neighbors import Make_NListNaive, NeighborListSet
    for rc, sep in ((4.6, 10.0), (3.1, 10.0), (15.0, 40.0)):
        ds = _ulp_distances(rc)
        x = np.zeros((2 * len(ds), 3))
        for k, d in enumerate(ds):
            x[2 * k] = (0.0, sep * k, 0.0)
            x[2 * k + 1] = (d, sep * k, 0.0)
        keep = [bool(np.sqrt(d * d) + 1e-13 < rc) for d in ds]
        assert any(keep) and not all(keep)        # the cases straddle the cutoff
        rows = Make_NListNaive(x, rc, len(x), 1)
        ref = O.nlist_naive(x, rc, len(x), 1)
        for i in range(len(x)):
            assert rows[i] == [int(j) for j in ref[i]]
        for k, kp in enumerate(keep):
            assert (2 * k + 1 in rows[2 * k]) == kp
        half = NeighborListSet(x[None], np.array([len(x)]), False, False).buildPairs(rc)
        np.testing.assert_array_equal(half.astype(np.int64),
                                      O.half_pairs(x[None], np.array([len(x)]), rc))


def test_cutoff_ulp_pairs_periodic_wrap(cuda):
    """The same straddling separations across the faces of a periodic box,
    through the engine's cell-list build (minimum image in float64)."""
    from tensormol_amd.engine import BPEEEngine
    from tests.test_gpu_parity import gpu_pairs_triples
    L = 31.06
    eng = BPEEEngine(H.model_for("water"), precision="fp32")
    for rc in (4.6, 3.1):
        ds = _ulp_distances(rc, 4)
        x, Z = [], []
        for k, d in enumerate(ds):
            # pairs on a 4 x 4 grid in (y, z), 7.765 A apart (more than 4.6 A, also
            # across the y and z faces), each pair split across the x face
            y = 0.5 + (L / 4) * (k % 4)
            z = 0.5 + (L / 4) * (k // 4)
            a = 0.25 + 1e-3 * k
            x += [(a, y, z), (a - d + L, y, z)]
            Z += [8, 1]
        x = np.array(x)
        Z = np.array(Z)
        ms = np.array([0, len(Z)])
        box = np.array([[L, L, L]])
        ws = eng.prepare(Z, ms, box, pos0=x)
        pos = torch.tensor(x, dtype=torch.float64, device=cuda)
        eng.build_lists(ws, pos)
        eng.check_flags(ws)
        trp, _, _ = gpu_pairs_triples(eng, ws)
        ref = O.nlist_naive(x, 4.6, None, 1, box[0])
        got = [sorted(int(j) for j in trp[trp[:, 1] == i, 2]) for i in range(len(Z))]
        assert got == [[int(j) for j in r] for r in ref]
        sub = O.nlist_naive(x, 3.1, None, 1, box[0])
        rows = sub if rc == 3.1 else ref
        keep = [2 * k + 1 in rows[2 * k] for k in range(len(ds))]
        assert any(keep) and not all(keep)          # the cases straddle the cutoff
        assert all(len(r) <= 1 for r in ref)        # no other pair within 4.6 A
        # the 3.1 A sub-list: which row entries carry the angular bit
        acnt = ws.acnt.cpu().numpy()
        np.testing.assert_array_equal(acnt[:len(Z)], [len(r) for r in sub])


# ------------------------------------------------------------ manager API
def _manager(modelname, precision="fp32"):
    from tensormol_amd.manage import TFMolManage, BPEEInstance
    mgr = TFMolManage("", None, False,
                      "fc_sqdiff_BP_Direct_EE_ChargeEncode_Update_vdw_DSF_elu_Normalize_Dropout",
                      False, False, model=H.model_for(modelname))
    mgr.Instances = BPEEInstance(mgr.model, precision=precision)
    return mgr


def test_eval_single_tuple_units_values(cuda):
    from tensormol_amd.containers import Mol
    from tensormol_amd.params import JOULEPERHARTREE
    mols = H.golden_mols("morphine")
    m = Mol(mols[0].atoms, mols[0].coords)
    mgr = _manager("chno")
    out = mgr.EvalBPDirectEEUpdateSingle(m, 4.6, 3.1, 15.0, True)
    assert len(out) == 8
    Etotal, Ebp, Ebp_atom, Ecc, Evdw, dip, q, F = out
    n = m.NAtoms()
    assert Etotal.shape == (1,) and Ebp_atom.shape == (1, n) and dip.shape == (1, 3)
    assert q.shape == (1, n) and F.shape == (1, n, 3)
    ref = H.oracle_eval("morphine", "chno")
    assert abs(Etotal[0] - ref["Etotal"][0]) < FP32_E_TOL * n
    assert abs(Ecc[0] - ref["Ecc"][0]) < FP32_E_TOL * n
    assert abs(Evdw[0] - ref["Evdw"][0]) < 1e-9 * n
    np.testing.assert_allclose(Etotal, Ebp + Ecc + Evdw, rtol=1e-12)
    np.testing.assert_allclose(q[0], ref["charge"], atol=1e-5)
    np.testing.assert_allclose(dip[0], ref["dipole"][0], atol=1e-4)
    # forces in J/mol/A: -JOULEPERHARTREE * dE/dx
    frms = np.sqrt(((F[0] / -JOULEPERHARTREE - ref["gradient"]) ** 2).mean())
    assert frms < FP32_G_TOL, frms
    # the vdW instance returns eight values; asking for six fails as in the reference
    with pytest.raises(ValueError):
        mgr.EvalBPDirectEEUpdateSingle(m, 4.6, 3.1, 15.0, False)
    with pytest.raises(ValueError):
        mgr.EvalBPDirectEEUpdateSingle(m, 4.6, 3.0, 15.0, True)     # cutoffs differ from the model
    qq = mgr.EvalBPDirectChargeSingle(m, 4.6, 3.1, 15.0)
    np.testing.assert_allclose(qq, q)


def test_eval_six_tuple_without_vdw(cuda):
    """A model without the dispersion term: HasVdw=False gives the six-tuple
    (Etotal, Ebp, Ecc, dipole, charges, force) and Etotal = Ebp + Ecc."""
    from tensormol_amd.containers import Mol
    from tensormol_amd.manage import TFMolManage, BPEEInstance
    from tensormol_amd.model import BPEEModel
    base = H.model_for("water")
    model = BPEEModel(base.eles, base.hidden, base.alpha, base.dsf_alpha, has_vdw=False,
                      energy_nets=base.energy_nets, charge_nets=base.charge_nets)
    mgr = TFMolManage("", None, False,
                      "fc_sqdiff_BP_Direct_EE_ChargeEncode_Update_vdw_DSF_elu", model=model)
    mgr.Instances = BPEEInstance(model, precision="fp32")
    Z, x = H.system("water10")[:2]
    out = mgr.EvalBPDirectEEUpdateSingle(Mol(Z, x), 4.6, 3.1, 15.0)
    assert len(out) == 6
    Etotal, Ebp, Ecc, dip, q, F = out
    sp = O.SymParams(model.eles)
    ref = O.evaluate(x, Z, np.array([0, len(Z)]), sp, model.energy_nets, model.charge_nets,
                     alpha=model.alpha, dsf_alpha=model.dsf_alpha, has_vdw=False)
    assert abs(Etotal[0] - ref["Etotal"][0]) < FP32_E_TOL * len(Z)
    np.testing.assert_allclose(Etotal, Ebp + Ecc, rtol=1e-12)
    assert F.shape == (1, len(Z), 3)


def test_eval_set_and_instance_protocol(cuda):
    from tensormol_amd.containers import MSet, Mol
    from tensormol_amd.params import JOULEPERHARTREE
    Z, x, ms, _ = H.system("chains")
    s = MSet("chains")
    for m in range(len(ms) - 1):
        s.mols.append(Mol(Z[ms[m]:ms[m + 1]], x[ms[m]:ms[m + 1]]))
    mgr = _manager("chno")
    Etotal, Ebp, Ebp_atom, Ecc, Evdw, dip, q, F = mgr.EvalBPDirectEEUpdateSet(s, 4.6, 3.1, 15.0, True)
    ref = H.oracle_eval("chains", "chno")
    natom = np.diff(ms)
    assert Etotal.shape == (len(natom),) and F.shape == (len(natom), natom.max(), 3)
    np.testing.assert_array_less(np.abs(Etotal - ref["Etotal"]), FP32_E_TOL * natom)
    for m in range(len(natom)):
        g = F[m, :natom[m]] / -JOULEPERHARTREE
        assert np.sqrt(((g - ref["gradient"][ms[m]:ms[m + 1]]) ** 2).mean()) < FP32_G_TOL
        assert np.all(F[m, natom[m]:] == 0.0)      # padding rows stay zero
    # the positional protocol directly: list inputs are accepted and rebuilt
    xyzs, Zs, nat, _ = _padded("chains")
    batch = [xyzs, Zs, np.zeros(len(nat)), np.zeros((len(nat), 3)), np.zeros_like(xyzs),
             "ignored", "ignored", "ignored", "ignored", 1.0 / nat]
    out = mgr.Instances.evaluate(batch)
    assert len(out) == 8 and isinstance(out[7], list)
    np.testing.assert_allclose(out[0], Etotal, rtol=1e-12)
    bad = list(batch)
    bad[2] = np.full(len(nat), np.nan)
    with pytest.raises(Exception, match="DontEatShit"):
        mgr.Instances.evaluate(bad)


def test_eval_single_periodic_from_tessellation(cuda):
    """The reference's periodic call: a tessellated molecule (real atoms first,
    then the 26 images of Lattice.TessLattice, Periodic.py:131-168) and nreal,
    no lattice argument.  The lattice is recovered from the images and the
    real atoms are evaluated under the minimum image."""
    from tensormol_amd.containers import Mol
    from tensormol_amd.params import JOULEPERHARTREE
    from tensormol_amd.periodic import Lattice
    Z, x, ms, boxes = H.system("waterbox10")
    L = boxes[0]
    lat = Lattice(np.diag(L))
    zt, xt = lat.TessLattice(Z.astype(np.uint8), x, 15.0)
    assert len(zt) == 27 * len(Z)
    mgr = _manager("water")
    n = len(Z)
    E, F = mgr.EvalBPDirectEEUpdateSinglePeriodic(Mol(zt, xt), 4.6, 3.1, 15.0, n, True)
    assert F.shape == (1, n, 3)
    ref = H.oracle_eval("waterbox10", "water")
    assert abs(E[0] - ref["Etotal"][0]) < FP32_E_TOL * n
    frms = np.sqrt(((F[0] / -JOULEPERHARTREE - ref["gradient"]) ** 2).mean())
    assert frms < FP32_G_TOL
    E2 = mgr.EvalBPDirectEEUpdateSinglePeriodic(Mol(zt, xt), 4.6, 3.1, 15.0, n, True, False)
    np.testing.assert_allclose(E2, E, rtol=1e-12)
    E3, F3, q3 = mgr.EvalBPDirectEEUpdateSinglePeriodic(Mol(zt, xt), 4.6, 3.1, 15.0, n, True, True,
                                                        True)
    assert q3.shape == (1, n) and abs(q3.sum()) < 1e-6


# ------------------------------------------------------------ host drivers on the GPU
def test_velocity_verlet_prop_with_manager(cuda, monkeypatch, tmp_path):
    from tensormol_amd import md as MD
    from tensormol_amd.containers import Mol
    from tensormol_amd.params import PARAMS, atomic_masses
    monkeypatch.setitem(PARAMS, "MDMaxStep", 8)
    monkeypatch.setitem(PARAMS, "MDdt", 0.5)
    monkeypatch.setitem(PARAMS, "MDV0", None)
    monkeypatch.setitem(PARAMS, "MDThermostat", None)
    monkeypatch.setitem(PARAMS, "results_dir", str(tmp_path) + "/")
    Z, x = H.system("water10")[:2]
    mgr = _manager("water")
    mol = Mol(Z, x)

    def EnAndForce(x_):
        mol.coords = x_
        out = mgr.EvalBPDirectEEUpdateSingle(mol, 4.6, 3.1, 15.0, True)
        return out[0][0], out[7][0]
    model = H.model_for("water")
    sp = O.SymParams(model.eles)

    def oracle_ef(x_):
        r = O.evaluate(x_, Z, np.array([0, len(Z)]), sp, model.energy_nets, model.charge_nets,
                       alpha=model.alpha, dsf_alpha=model.dsf_alpha)
        return r["Etotal"][0], -O.JOULEPERHARTREE * r["gradient"]
    vv = MD.VelocityVerlet(None, Mol(Z, x), "w", EnAndForce)
    vv.Prop()
    m = atomic_masses(Z)
    xo, vo, ao = x.copy(), np.zeros_like(x), np.zeros_like(x)
    for _ in range(8):
        xo, vo, ao, eo = O.vv_step(xo, vo, ao, m, 0.5, oracle_ef)
    assert np.abs(vv.x - xo).max() < 1e-5     # FP32 forces over 8 steps
    assert abs(vv.md_log[-1, 5] - eo) < FP32_E_TOL * len(Z)


def test_geom_optimizer_with_manager(cuda, monkeypatch, tmp_path):
    from tensormol_amd.containers import Mol
    from tensormol_amd.opt import GeomOptimizer
    from tensormol_amd.params import PARAMS
    monkeypatch.setitem(PARAMS, "results_dir", str(tmp_path) + "/")
    monkeypatch.setitem(PARAMS, "OptMaxCycles", 6)
    Z, x = H.system("water10")[:2]
    mgr = _manager("water")
    mol = Mol(Z, x)

    def ef(x_, DoForce=True):
        mol.coords = x_
        out = mgr.EvalBPDirectEEUpdateSingle(mol, 4.6, 3.1, 15.0, True)
        return (out[0][0], out[7][0]) if DoForce else out[0][0]
    e0 = ef(x, False)
    out = GeomOptimizer(ef).Opt(Mol(Z, x), filename="w")
    assert out.properties["Energy"] < e0
    log = open(tmp_path / "w.xyz").read()
    es = [float(l.split(";;;Energy ")[1].split(";;;")[0]) for l in log.splitlines() if ";;;Energy " in l]
    assert len(es) == out.properties["Step"] + 1
    assert all(b <= a + 1e-9 for a, b in zip(es, es[1:]))


# ------------------------------------------------------------ periodic drivers
def _periodic_setup():
    from tensormol_amd.containers import Mol
    from tensormol_amd.periodic import PeriodicForce
    Z, x, ms, boxes = H.system("waterbox10")
    L = boxes[0]
    mgr = _manager("water")

    def EnAndForce(z_, x_, nreal_, DoForce=True):
        """The reference sample's local force (samples/test_h2o.py:1852-1862)."""
        mtmp = Mol(z_, x_)
        if DoForce:
            en, f = mgr.EvalBPDirectEEUpdateSinglePeriodic(mtmp, 4.6, 3.1, 15.0, nreal_, True, DoForce)
            return en[0], f[0]
        en = mgr.EvalBPDirectEEUpdateSinglePeriodic(mtmp, 4.6, 3.1, 15.0, nreal_, True, DoForce)
        return en[0]
    PF = PeriodicForce(Mol(Z, x), np.diag(L))
    PF.BindForce(EnAndForce, 15.0)
    model = H.model_for("water")
    sp = O.SymParams(model.eles)

    def oracle_ef(xx):
        r = O.evaluate(xx, Z, np.array([0, len(Z)]), sp, model.energy_nets, model.charge_nets,
                       alpha=model.alpha, dsf_alpha=model.dsf_alpha, boxes=boxes)
        return r["Etotal"][0], -O.JOULEPERHARTREE * r["gradient"]
    return Z, L, PF, oracle_ef


def test_periodic_force_vs_oracle(cuda):
    from tensormol_amd.params import JOULEPERHARTREE
    Z, L, PF, oracle_ef = _periodic_setup()
    x0 = PF.mol0.coords
    assert np.all((x0 >= 0) & (x0 < L))                 # centred and wrapped into the cell
    E, F = PF(x0)
    e_ref, f_ref = oracle_ef(x0)
    assert abs(E - e_ref) < FP32_E_TOL * len(Z)
    assert np.sqrt((((F - f_ref) / JOULEPERHARTREE) ** 2).mean()) < FP32_G_TOL
    E2, _ = PF(x0 + L * np.array([1.0, -2.0, 0.0]))      # whole-cell shifts change nothing
    assert abs(E2 - E) < 1e-9 * abs(E)
    e_only, _ = PF(x0, DoForce=False)
    assert abs(e_only - E) < 1e-9 * abs(E)
    assert abs(PF.Density() - 0.99) < 0.05              # 1000 waters in a 31.06 A cube


@pytest.mark.parametrize("thermostat", [None, "Nose"])
def test_periodic_velocity_verlet_vs_oracle(cuda, monkeypatch, tmp_path, thermostat):
    """PeriodicVelocityVerlet (PeriodicMD.py:77-146) over the periodic manager
    call, against the oracle's minimum-image Velocity Verlet / Nose-Hoover steps
    from the same start (a = 0 and v = 0 on the first step, as the reference)."""
    from tensormol_amd.params import PARAMS, atomic_masses
    from tensormol_amd.periodic import PeriodicVelocityVerlet
    monkeypatch.setitem(PARAMS, "MDMaxStep", 4)
    monkeypatch.setitem(PARAMS, "MDdt", 0.5)
    monkeypatch.setitem(PARAMS, "MDV0", None)
    monkeypatch.setitem(PARAMS, "MDThermostat", thermostat)
    monkeypatch.setitem(PARAMS, "results_dir", str(tmp_path) + "/")
    Z, L, PF, oracle_ef = _periodic_setup()
    md = PeriodicVelocityVerlet(PF, "pw")
    x0 = md.x.copy()
    md.Prop()
    m = atomic_masses(Z)
    x, v, a, eta = x0.copy(), np.zeros_like(x0), np.zeros_like(x0), 0.0
    for _ in range(4):
        if thermostat is None:
            x, v, a, e = O.vv_step(x, v, a, m, 0.5, oracle_ef)
        else:
            x, v, a, e, eta = O.nose_step(x, v, a, m, 0.5, oracle_ef, eta, PARAMS["MDTemp"], 0.5)
    d = md.x - x
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-6
    assert np.all((md.x >= 0) & (md.x < L))
    assert abs(md.md_log[-1, 5] - e) < FP32_E_TOL * len(Z)
    if thermostat == "Nose":
        assert abs(md.Tstat.eta - eta) < 1e-3 * abs(eta) and eta != 0.0


def test_device_fire_vs_oracle(cuda):
    """DeviceFIRE (the serving optimiser) on the FP32 path against the oracle's
    FIRE step (Bitzek et al. 2006) with the same parameters, 20 iterations on
    the gentle water model, and one batch of two molecules (each with its own
    time step and mixing state)."""
    from tensormol_amd.engine import BPEEEngine
    from tensormol_amd.opt import DeviceFIRE
    from tensormol_amd.params import atomic_masses
    from tests.test_gpu_md import _gentle
    Z1, x1, _, model, ef1 = _gentle()
    Z2, x2 = Z1[:12], x1[:12] + np.array([40.0, 0.0, 0.0])
    sp = O.SymParams(model.eles)

    def ef(Z, xx):
        r = O.evaluate(xx, Z, np.array([0, len(Z)]), sp, model.energy_nets, model.charge_nets,
                       alpha=model.alpha, dsf_alpha=model.dsf_alpha)
        return -O.JOULEPERHARTREE * r["gradient"]
    Z = np.concatenate([Z1, Z2])
    ms = np.array([0, len(Z1), len(Z)])
    eng = BPEEEngine(model, precision="fp32")
    ws = eng.prepare(Z, ms, None, pos0=np.concatenate([x1, x2]))
    fire = DeviceFIRE(eng, ws, Z, np.concatenate([x1, x2]), dt0=1.0, fmax=1e-9, maxstep=0.1)
    states = [[1.0, 0.1, 0], [1.0, 0.1, 0]]
    xs = [x1.copy(), x2.copy()]
    vs = [np.zeros_like(x1), np.zeros_like(x2)]
    for _ in range(20):
        fire.step()
        for k, Zk in enumerate((Z1, Z2)):
            xs[k], vs[k], states[k] = O.fire_step(xs[k], vs[k], ef(Zk, xs[k]), atomic_masses(Zk),
                                                  states[k], dtmax=10.0, maxstep=0.1)
    torch.cuda.synchronize()
    xg = fire.pos.cpu().numpy()
    st = fire.state.cpu().numpy()
    for k in range(2):
        dx = np.abs(xg[ms[k]:ms[k + 1]] - xs[k]).max()
        print(f"FIRE molecule {k}: |dx| {dx:.2e} A, dt gpu {st[k, 0]:.4f} oracle {states[k][0]:.4f}")
        assert dx < 1e-5
        assert abs(st[k, 0] - states[k][0]) < 1e-12 and st[k, 2] == states[k][2]


def test_manager_evaluates_reference_checkpoint(cuda, tmp_path):
    """A network saved by the reference as a TF checkpoint (<networks_directory>/<Name>/
    <Name>-chk-<step>, variable names of TFMolInstanceDirect.py:5179-5264) is found by
    TFMolManage(Name) and evaluates as the same weights do: bit-identical to a manager
    holding the source model, and within the FP32 gates of the oracle."""
    from tensormol_amd import tfbundle as TB
    from tensormol_amd.containers import Mol
    from tensormol_amd.manage import TFMolManage, BPEEInstance
    from tensormol_amd.params import PARAMS
    from tests import tfbundle_writer as TW
    src = H.model_for("water")
    tensors = {}
    for Zel in src.eles:
        for nets, names in zip((src.energy_nets, src.charge_nets),
                               TB.bpee_variable_names(Zel, len(src.hidden))):
            for (Wm, b), (wn, bn) in zip(nets[Zel], names):
                tensors[wn], tensors[bn] = Wm, b
                tensors[wn + "/Adam"] = np.zeros_like(Wm)
    d = tmp_path / "water_ref"
    d.mkdir()
    TW.write_bundle(str(d / "water_ref-chk-300"), tensors, block_entries=7)
    nettype = "fc_sqdiff_BP_Direct_EE_ChargeEncode_Update_vdw_DSF_elu_Normalize_Dropout"
    saved = {k: PARAMS[k] for k in ("networks_directory", "NeuronType", "DSFAlpha")}
    PARAMS.update(networks_directory=str(tmp_path), NeuronType="softplus", DSFAlpha=src.dsf_alpha)
    try:
        mgr = TFMolManage("water_ref", None, False, nettype, False, False)
    finally:
        PARAMS.update(saved)
    assert mgr.model.eles == src.eles and mgr.model.hidden == src.hidden
    mgr.Instances = BPEEInstance(mgr.model, precision="fp32")
    ref_mgr = TFMolManage("", None, False, nettype, False, False, model=src)
    ref_mgr.Instances = BPEEInstance(src, precision="fp32")
    Z, x = H.system("water10")[:2]
    out = mgr.EvalBPDirectEEUpdateSingle(Mol(Z, x), 4.6, 3.1, 15.0, True)
    same = ref_mgr.EvalBPDirectEEUpdateSingle(Mol(Z, x), 4.6, 3.1, 15.0, True)
    for a, b in zip(out, same):
        np.testing.assert_array_equal(a, b)
    sp = O.SymParams(src.eles)
    ref = O.evaluate(x, Z, np.array([0, len(Z)]), sp, src.energy_nets, src.charge_nets,
                     alpha=src.alpha, dsf_alpha=src.dsf_alpha, has_vdw=src.has_vdw)
    assert abs(out[0][0] - ref["Etotal"][0]) < FP32_E_TOL * len(Z)
