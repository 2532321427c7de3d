"""BP+EE model definition: element list, AN1 descriptor grid and the per-element
energy and charge networks.

Mirrors the state of MolInstance_DirectBP_EE_ChargeEncode_Update_vdw_DSF_elu_
Normalize_Dropout (TFMolInstanceDirect.py:4965-6010 and its parents): eles
sorted ascending (TensorMolData.py:1231-1232), eles_pairs with a <= b
(TFMolInstanceDirect.py:1263-1267), descriptor length
nele*32 + nelep*64 (:3158), per-element dense networks with HiddenLayers
widths and a linear regression output (:5164-5212, :5219-5272).

Weights live here as float64 numpy master copies in the repository's own
format (one .npz); the GPU engine packs them into MFMA tile order.
"""
import math

import numpy as np

from .params import PARAMS, BOHRPERA, JOULEPERHARTREE, C6_COFF, ATOMIC_VDW_RADIUS, an1_grids


class SymConfig:
    """AN1 symmetry-function grid (TMParams.py:26-38, TFMolInstanceDirect.py:3128-3164)."""

    def __init__(self, eles, p=PARAMS):
        self.eles = [int(e) for e in sorted(eles)]
        self.nele = len(self.eles)
        self.epairs = [(self.eles[a], self.eles[b]) for a in range(self.nele) for b in range(a, self.nele)]
        self.nelep = len(self.epairs)
        self.rr = float(p["AN1_r_Rc"])
        self.ra = float(p["AN1_a_Rc"])
        self.eta = float(p["AN1_eta"])
        self.zeta = float(p["AN1_zeta"])
        self.nr_r = int(p["AN1_num_r_Rs"])
        self.nr_a = int(p["AN1_num_a_Rs"])
        self.nth_a = int(p["AN1_num_a_As"])
        self.rs_r, self.rs_a, self.thetas = an1_grids(p)
        self.D = self.nele * self.nr_r + self.nelep * self.nr_a * self.nth_a

    def pair_index(self, a, b):
        """Index of the unordered element pair (element indices a, b)."""
        a, b = min(a, b), max(a, b)
        return a * self.nele - a * (a - 1) // 2 + (b - a)


def glorot_layers(rng, sizes):
    """Glorot-uniform weights, zero biases."""
    out = []
    for fi, fo in zip(sizes[:-1], sizes[1:]):
        lim = math.sqrt(6.0 / (fi + fo))
        out.append((rng.uniform(-lim, lim, size=(fi, fo)), np.zeros(fo)))
    return out


def truncnorm_layers(rng, sizes):
    """Truncated-normal weights with stddev 1/(10+sqrt(fan_in)) and zero biases,
    the reference initialiser (TFMolInstanceDirect.py:5179-5201, TFInstance.py:277-297)."""
    out = []
    for fi, fo in zip(sizes[:-1], sizes[1:]):
        sd = 1.0 / (10.0 + math.sqrt(float(fi)))
        w = rng.normal(0.0, sd, size=(fi, fo))
        bad = np.abs(w) > 2 * sd
        while bad.any():
            w[bad] = rng.normal(0.0, sd, size=int(bad.sum()))
            bad = np.abs(w) > 2 * sd
        out.append((w, np.zeros(fo)))
    return out


class BPEEModel:
    """Parameters of one BP+EE potential.

    activation_alpha: the softplus sharpness (1 = tf.nn.softplus; 100 =
    sigmoid_with_param as used by the shipped TensorMol-0.1 networks).
    """

    def __init__(self, eles, hidden=None, activation_alpha=1.0, dsf_alpha=None, has_vdw=True,
                 seed=0, init="glorot", p=PARAMS, energy_nets=None, charge_nets=None):
        self.sym = SymConfig(eles, p)
        self.eles = self.sym.eles
        self.hidden = list(hidden if hidden is not None else p["HiddenLayers"])
        self.alpha = float(activation_alpha)
        self.dsf_alpha = float(p["DSFAlpha"] if dsf_alpha is None else dsf_alpha)
        self.has_vdw = bool(has_vdw)
        self.elu_width = float(p["Elu_Width"])
        self.ee_off = float(p["EECutoffOff"])
        self.poly_width = float(p["Poly_Width"])
        sizes = [self.sym.D] + self.hidden + [1]
        if energy_nets is None:
            rng = np.random.default_rng(seed)
            mk = glorot_layers if init == "glorot" else truncnorm_layers
            energy_nets = {e: mk(rng, sizes) for e in self.eles}
            charge_nets = {e: mk(rng, sizes) for e in self.eles}
        self.energy_nets = energy_nets
        self.charge_nets = charge_nets

    # vdW constants in atomic units (TFMolInstanceDirect.py:3764-3767)
    def vdw_tables(self):
        c6 = np.array([C6_COFF[e] * (BOHRPERA * 10.0) ** 6.0 / JOULEPERHARTREE for e in self.eles])
        r0 = np.array([ATOMIC_VDW_RADIUS[e] * BOHRPERA for e in self.eles])
        return c6, r0

    def element_index(self, Z):
        """Index of each atom's element in self.eles (int32)."""
        Z = np.asarray(Z, dtype=np.int64)
        lut = np.full(max(max(self.eles), int(Z.max()) if Z.size else 0) + 1, -1, dtype=np.int32)
        lut[np.array(self.eles)] = np.arange(len(self.eles), dtype=np.int32)
        if Z.size and Z.min() < 0:
            raise ValueError(f"element {int(Z.min())} not in model elements {self.eles}")
        idx = lut[Z]
        if Z.size and idx.min() < 0:
            raise ValueError(f"element {int(Z[np.argmin(idx)])} not in model elements {self.eles}")
        return idx

    def save(self, path):
        arrs = {"eles": np.array(self.eles), "hidden": np.array(self.hidden),
                "scalars": np.array([self.alpha, self.dsf_alpha, float(self.has_vdw)])}
        for tag, nets in (("E", self.energy_nets), ("Q", self.charge_nets)):
            for e, layers in nets.items():
                for k, (W, b) in enumerate(layers):
                    arrs[f"{tag}_{e}_{k}_W"] = W
                    arrs[f"{tag}_{e}_{k}_b"] = b
        np.savez(path, **arrs)

    @classmethod
    def load(cls, path):
        f = np.load(path, allow_pickle=False)
        eles = [int(e) for e in f["eles"]]
        hidden = [int(h) for h in f["hidden"]]
        alpha, dsf, vdw = [float(v) for v in f["scalars"]]
        nl = len(hidden) + 1
        nets = {}
        for tag in ("E", "Q"):
            nets[tag] = {e: [(f[f"{tag}_{e}_{k}_W"], f[f"{tag}_{e}_{k}_b"]) for k in range(nl)] for e in eles}
        return cls(eles, hidden, alpha, dsf, bool(vdw), energy_nets=nets["E"], charge_nets=nets["Q"])
