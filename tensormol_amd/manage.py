"""Drop-in manager API: TFMolManage and the BP-EE model instance.

Mirrors TensorMol/TFNetworks/TFMolManage.py (constructor :38-60, the BP-EE
evaluation entry points :1260-1465) and the instance protocol of
MolInstance_DirectBP_EE_ChargeEncode_Update_vdw_DSF_elu_Normalize_Dropout
(evaluate :5684-5711, evaluate_periodic :5919-5947): same method names,
argument meaning, return tuples, shapes and units (A in; Ha, e, e*bohr and
J/mol/A out).  The list arguments of the reference's positional batch
protocol are accepted and ignored: the engine rebuilds them on the device.

Weights are the repository's own format (BPEEModel.save/load) found as
<networks_directory>/<Name>.npz, else the last TF checkpoint the reference saved under
<networks_directory>/<Name>/ (weights only); an unknown name gets seeded weights.
"""
import os

import numpy as np
import torch

from .containers import Mol, MSet
from .engine import BPEEEngine
from .model import BPEEModel
from .params import PARAMS, JOULEPERHARTREE

BPEE_NETTYPES = (
    "fc_sqdiff_BP_Direct_EE_ChargeEncode_Update_vdw_DSF_elu_Normalize_Dropout",
    "fc_sqdiff_BP_Direct_EE_ChargeEncode_Update_vdw_DSF_elu_Normalize",
    "fc_sqdiff_BP_Direct_EE_ChargeEncode_Update_vdw_DSF_elu",
)


class MolDigester:
    """Descriptor choice of the reference (DigestMol.py:11); only ANI1_Sym_Direct
    (the AN1 radial + angular set) is supported."""

    def __init__(self, eles_, name_="ANI1_Sym_Direct", OType_="EnergyAndDipole", SensedAtoms_=None):
        if name_ != "ANI1_Sym_Direct":
            raise NotImplementedError(f"digester {name_!r}: only ANI1_Sym_Direct is built")
        self.eles = [int(e) for e in np.asarray(eles_).reshape(-1)]
        self.name = name_
        self.OType = OType_


class TensorMolData_BP_Direct_EE_WithEle:
    """Training-set wrapper of TensorMolData.py:1849-1904: carries the set, the
    sorted element list and MaxNAtoms."""

    def __init__(self, MSet_=None, Dig_=None, Name_=None, order_=3, num_indis_=1, type_="mol",
                 WithGrad_=False):
        self.set = MSet_
        self.dig = Dig_
        self.name = Name_ if Name_ is not None else (MSet_.name if MSet_ is not None else "set")
        self.eles = sorted(int(e) for e in (MSet_.AtomTypes() if MSet_ is not None else Dig_.eles))
        if Dig_ is not None:
            self.eles = sorted(set(self.eles) | set(Dig_.eles)) if MSet_ is None else self.eles
        self.MaxNAtoms = MSet_.MaxNAtoms() if MSet_ is not None else 0
        self.HasGrad = WithGrad_


def _activation_alpha():
    nt = PARAMS["NeuronType"]
    if nt == "sigmoid_with_param":
        return float(PARAMS["sigmoid_alpha"])
    if nt == "softplus":
        return 1.0
    raise NotImplementedError(f"NeuronType {nt!r}: the engine implements softplus and "
                              "sigmoid_with_param")


class BPEEInstance:
    """Evaluation side of the BP-EE instance (TFMolInstanceDirect.py:4965-5772)."""

    def __init__(self, model, precision=None):
        self.model = model
        self.eles = model.eles
        self.n_eles = len(self.eles)
        self.eles_np = np.asarray(self.eles).reshape((self.n_eles, 1))
        self.eles_pairs = [[a, b] for (a, b) in model.sym.epairs]
        self.eles_pairs_np = np.asarray(self.eles_pairs)
        self.precision = precision or PARAMS["MLPPrecision"]
        self.engine = None
        self._ws_key = None
        self._ws = None

    def _engine(self):
        if self.engine is None:
            self.engine = BPEEEngine(self.model, precision=self.precision)
        return self.engine

    def _workspace(self, Z, ms, boxes, pos0):
        key = (tuple(int(z) for z in Z), tuple(int(v) for v in ms),
               None if boxes is None else tuple(np.asarray(boxes).ravel()))
        if key != self._ws_key:
            self._ws = self._engine().prepare(Z, ms, boxes, pos0=pos0)
            self._ws_key = key
        return self._ws

    def _run(self, xyzs, Zs, natom, boxes=None, forces=True):
        B = xyzs.shape[0]
        natom = np.asarray(natom, dtype=np.int64)
        ms = np.concatenate([[0], np.cumsum(natom)])
        Z = np.concatenate([Zs[m, :natom[m]] for m in range(B)]).astype(np.int64)
        x = np.concatenate([xyzs[m, :natom[m]] for m in range(B)]).astype(np.float64)
        eng = self._engine()
        ws = self._workspace(Z, ms, boxes, x)
        pos = torch.tensor(x, dtype=torch.float64, device=eng.device)
        eng.run(ws, pos, forces=forces)
        torch.cuda.synchronize(eng.device)
        r = eng.results(ws)
        maxn = xyzs.shape[1]
        pad = lambda v: np.stack([np.concatenate([v[ms[m]:ms[m + 1]],
                                                  np.zeros((maxn - natom[m],) + v.shape[1:])])
                                  for m in range(B)])
        return r, pad

    def evaluate(self, batch_data):
        """batch_data = [xyzs (B,MaxN,3), Zs (B,MaxN), E*, mu*, grad*, rad_p_ele,
        ang_t_elep, rad_eep, mil_jk, 1/natom] -> (Etotal, Ebp, Ebp_atom, Ecc,
        Evdw, mol_dipole, atom_charge, gradient) with gradient = [dE/dx (B,MaxN,3)]
        in Ha/A (TFMolInstanceDirect.py:5684-5711)."""
        xyzs, Zs = np.asarray(batch_data[0]), np.asarray(batch_data[1])
        if not np.all(np.isfinite(batch_data[2])):
            raise Exception("DontEatShit")   # TFMolInstanceDirect.py:5158-5160
        natom = np.rint(1.0 / np.asarray(batch_data[9])).astype(np.int64)
        r, pad = self._run(xyzs, Zs, natom)
        return (r["Etotal"], r["Ebp"], pad(r["Ebp_atom"]), r["Ecc"], r["Evdw"], r["dipole"],
                pad(r["charge"]), [pad(r["gradient"])])

    def evaluate_periodic(self, batch_data, nreal, DoForce=True, lattice=None):
        """Periodic variant on the first nreal atoms under the minimum-image
        convention of an orthorhombic lattice (edges (3,) or a diagonal (3,3))."""
        if lattice is None:
            raise ValueError("evaluate_periodic needs the lattice (box edges in A)")
        L = np.asarray(lattice, dtype=np.float64)
        edges = np.diag(L) if L.ndim == 2 else L.reshape(3)
        xyzs = np.asarray(batch_data[0])[:, :nreal]
        Zs = np.asarray(batch_data[1])[:, :nreal]
        r, pad = self._run(xyzs, Zs, np.array([nreal]), boxes=edges[None], forces=DoForce)
        if not DoForce:
            return r["Etotal"]
        return (r["Etotal"], r["Ebp"], pad(r["Ebp_atom"]), r["Ecc"], r["Evdw"], r["dipole"],
                pad(r["charge"]), [pad(r["gradient"])])


class TFMolManage:
    """TFMolManage(Name_, TData_, Train_, NetType_, RandomTData_, Trainable_) (TFMolManage.py:38)."""

    def __init__(self, Name_="", TData_=None, Train_=False, NetType_="fc_sqdiff", RandomTData_=True,
                 Trainable_=True, model=None):
        if NetType_ not in BPEE_NETTYPES:
            raise NotImplementedError(f"NetType {NetType_!r}: this engine implements the BP-EE types "
                                      f"{BPEE_NETTYPES}")
        if Train_:
            raise NotImplementedError("training through TFMolManage is not part of this engine")
        self.name = Name_
        self.NetType = NetType_
        self.TData = TData_
        path = os.path.join(PARAMS["networks_directory"], f"{Name_}.npz") if Name_ else None
        chk = None
        if model is None and Name_ and not (path and os.path.exists(path)):
            # a network trained by the reference: its last TF checkpoint
            # (TFInstance.py:94, 246-258), weights only
            from .tfbundle import find_last_checkpoint
            chk = find_last_checkpoint(os.path.join(PARAMS["networks_directory"], Name_), Name_)
        if model is None:
            if path and os.path.exists(path):
                model = BPEEModel.load(path)
            elif chk is not None:
                from .tfbundle import model_from_checkpoint
                model = model_from_checkpoint(chk, TData_.eles if TData_ is not None else None,
                                              activation_alpha=_activation_alpha(),
                                              dsf_alpha=PARAMS["DSFAlpha"])
            else:
                eles = TData_.eles if TData_ is not None else [1, 6, 7, 8]
                model = BPEEModel(eles, activation_alpha=_activation_alpha(),
                                  dsf_alpha=PARAMS["DSFAlpha"], seed=0)
        self.model = model
        self.Instances = BPEEInstance(model)

    def _check_cuts(self, Rr_cut, Ra_cut, Ree_cut):
        m = self.model
        if abs(Rr_cut - m.sym.rr) > 1e-12 or abs(Ra_cut - m.sym.ra) > 1e-12 or abs(Ree_cut - m.ee_off) > 1e-12:
            raise ValueError(f"cutoffs ({Rr_cut}, {Ra_cut}, {Ree_cut}) differ from the model's "
                             f"({m.sym.rr}, {m.sym.ra}, {m.ee_off})")

    @staticmethod
    def _pad(mols):
        n = len(mols)
        maxn = max(m.NAtoms() for m in mols)
        xyzs = np.zeros((n, maxn, 3))
        Zs = np.zeros((n, maxn), dtype=np.int32)
        natom = np.zeros(n, dtype=np.int32)
        for i, m in enumerate(mols):
            xyzs[i, :m.NAtoms()] = m.coords
            Zs[i, :m.NAtoms()] = m.atoms
            natom[i] = m.NAtoms()
        return xyzs, Zs, natom

    def _eval(self, mols, HasVdw):
        """The tuple follows HasVdw as in the reference: eight values
        (Etotal, Ebp, Ebp_atom, Ecc, Evdw, dipole, charges, force) with it, six
        (Etotal, Ebp, Ecc, dipole, charges, force) without.  The reference's
        vdW instance always returns eight, so HasVdw=False fails there when
        unpacking; here it raises ValueError for a model with the dispersion
        term and returns the six-tuple for a model without it."""
        if not HasVdw and self.model.has_vdw:
            raise ValueError("this BP-EE network includes the vdW term and returns eight values; "
                             "call with HasVdw=True (the reference fails unpacking six)")
        xyzs, Zs, natom = self._pad(mols)
        n = len(mols)
        batch = [xyzs, Zs, np.zeros(n), np.zeros((n, 3)), np.zeros_like(xyzs), None, None, None, None,
                 1.0 / natom]
        Etotal, Ebp, Ebp_atom, Ecc, Evdw, dip, q, grad = self.Instances.evaluate(batch)
        if not HasVdw:
            return Etotal, Ebp, Ecc, dip, q, -JOULEPERHARTREE * grad[0]
        return Etotal, Ebp, Ebp_atom, Ecc, Evdw, dip, q, -JOULEPERHARTREE * grad[0]

    def EvalBPDirectEEUpdateSingle(self, mol, Rr_cut, Ra_cut, Ree_cut, HasVdw=False):
        """TFMolManage.py:1290-1321."""
        self._check_cuts(Rr_cut, Ra_cut, Ree_cut)
        return self._eval([mol], HasVdw)

    def EvalBPDirectEEUpdateSet(self, mol_set, Rr_cut, Ra_cut, Ree_cut, HasVdw=False):
        """TFMolManage.py:1260-1288: all molecules of the set in one packed launch."""
        self._check_cuts(Rr_cut, Ra_cut, Ree_cut)
        return self._eval(mol_set.mols, HasVdw)

    def EvalBPDirectEELinearSingle(self, mol, Rr_cut, Ra_cut, Ree_cut, HasVdw=False):
        """TFMolManage.py:1410-1439 (the linear-scaling variant; same results)."""
        return self.EvalBPDirectEEUpdateSingle(mol, Rr_cut, Ra_cut, Ree_cut, HasVdw)

    def EvalBPDirectChargeSingle(self, mol, Rr_cut, Ra_cut, Ree_cut):
        """TFMolManage.py:1444-1465: atomic charges (1, N)."""
        self._check_cuts(Rr_cut, Ra_cut, Ree_cut)
        r = self._eval([mol], True)
        return r[6]

    def EvalBPDirectEEUpdateSinglePeriodic(self, mol, Rr_cut, Ra_cut, Ree_cut, nreal, HasVdw=True,
                                           DoForce=True, DoCharge=False, lattice=None):
        """TFMolManage.py:1323-1358: energy of the cell and forces on the first
        nreal atoms of mol.  As in the reference, mol is the tessellated
        molecule (real atoms first, then the images of Lattice.TessLattice);
        the lattice is read back from the images (or taken from `lattice` /
        mol.properties["Lattice"]) and the real atoms are evaluated with
        minimum-image lists.  Returns Etotal, force (1, nreal, 3) J/mol/A
        [, charges (1, nreal)], or Etotal alone without DoForce."""
        from .periodic import lattice_from_images, orthorhombic_edges
        self._check_cuts(Rr_cut, Ra_cut, Ree_cut)
        lat = lattice
        if lat is None:
            lat = mol.properties.get("Lattice", mol.properties.get("lattice"))
        if lat is None and mol.NAtoms() > nreal:
            lat = lattice_from_images(mol.coords, nreal)
        if lat is None:
            raise ValueError("no lattice: pass the tessellated molecule (real atoms, then the "
                             "images of Lattice.TessLattice), a Lattice property or lattice=")
        lat = orthorhombic_edges(lat)
        xyzs, Zs, natom = self._pad([mol])
        batch = [xyzs, Zs, np.zeros(1), np.zeros((1, 3)), np.zeros_like(xyzs), None, None, None, None,
                 np.array([1.0 / nreal])]
        if not DoForce:
            return self.Instances.evaluate_periodic(batch, nreal, False, lat)
        Etotal, Ebp, Ebp_atom, Ecc, Evdw, dip, q, grad = self.Instances.evaluate_periodic(
            batch, nreal, True, lat)
        force = -JOULEPERHARTREE * grad[0][0][:nreal].reshape(1, nreal, 3)
        if not DoCharge:
            return Etotal, force
        return Etotal, force, q[0][:nreal].reshape(1, nreal)
