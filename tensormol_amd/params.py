"""Physical constants and the PARAMS dictionary of the BP+EE path.

Values restate TensorMol's PhysicalData.py:24-54 and the TMParams.py defaults
that the BP+EE evaluation, MD and optimiser drivers read
(TMParams.py:26-38,57-58,66,79,97-120,140-154).  Drop-in scripts mutate PARAMS
before constructing objects, as with the reference; the engine copies what it
needs into a frozen config at construction and reads nothing global inside a
launch.
"""
import math

import numpy as np

BOHRPERA = 1.889725989
KJPERHARTREE = 2625.499638
JOULEPERHARTREE = KJPERHARTREE * 1000.0
KCALPERHARTREE = 627.509474
IDEALGASR = 8.3144621
FSPERAU = 0.0241888
AVOCONST = 6.02214086e23
PI_LITERAL = 3.14159265359
TWO_OVER_SQRTPI = 1.1283791671

ATOMICMASSESAMU = {1: 1.00794, 2: 4.002602, 3: 6.941, 4: 9.012182, 5: 10.811, 6: 12.0107,
                   7: 14.0067, 8: 15.9994, 9: 18.9984032, 10: 20.1791}
C6_COFF = {1: 0.14, 2: 0.08, 3: 1.16, 4: 1.61, 5: 3.13, 6: 1.75, 7: 1.23, 8: 0.70, 9: 0.75, 10: 0.63}
ATOMIC_VDW_RADIUS = {1: 1.001, 2: 1.012, 3: 0.825, 4: 1.408, 5: 1.485, 6: 1.452, 7: 1.397,
                     8: 1.342, 9: 1.287, 10: 1.243}


_MASS_LUT = np.full(max(ATOMICMASSESAMU) + 1, np.nan)
for _z, _m in ATOMICMASSESAMU.items():
    _MASS_LUT[_z] = 0.000999977 * _m


def atomic_masses(Z):
    """kg/mol per atom: 0.000999977 * amu (PhysicalData.py:54)."""
    Z = np.asarray(Z, dtype=np.int64)
    if Z.size and (Z.min() < 0 or Z.max() >= len(_MASS_LUT) or np.isnan(_MASS_LUT[Z]).any()):
        bad = sorted(set(int(z) for z in Z if z < 0 or z >= len(_MASS_LUT) or np.isnan(_MASS_LUT[z])))
        raise KeyError(bad[0])
    return _MASS_LUT[Z].astype(np.float64)


class _Params(dict):
    def __init__(self):
        super().__init__()
        self["AN1_r_Rc"] = 4.6
        self["AN1_a_Rc"] = 3.1
        self["AN1_eta"] = 4.0
        self["AN1_zeta"] = 8.0
        self["AN1_num_r_Rs"] = 32
        self["AN1_num_a_Rs"] = 8
        self["AN1_num_a_As"] = 8
        self["HiddenLayers"] = [512, 512, 512]
        self["NeuronType"] = "softplus"
        self["sigmoid_alpha"] = 100.0
        self["AddEcc"] = True
        self["Poly_Width"] = 4.6
        self["Elu_Width"] = 4.6
        self["EECutoffOn"] = 0.0
        self["EECutoffOff"] = 15.0
        self["DSFAlpha"] = 0.18
        self["KeepProb"] = [1.0, 1.0, 1.0, 1.0]
        self["MDdt"] = 0.2
        self["MDTemp"] = 300.0
        self["MDMaxStep"] = 20000
        self["MDThermostat"] = None
        self["MDV0"] = "Random"
        self["MDLogTrajectory"] = False
        self["OptMaxCycles"] = 50
        self["OptThresh"] = 0.0001
        self["OptMaxStep"] = 0.1
        self["OptStepSize"] = 0.1
        self["OptMomentum"] = 0.0
        self["OptMomentumDecay"] = 0.8
        self["GSSearchAlpha"] = 0.05
        self["OptLatticeStep"] = 0.050
        self["RemoveInvariant"] = True
        self["ConjGrad_Restart"] = 10
        self["FIREMaxStep"] = 0.1
        self["tm_root"] = "."
        self["sets_dir"] = "./datasets/"
        self["networks_directory"] = "./networks/"
        self["results_dir"] = "./results/"
        # precision of the MLP GEMMs on the GPU: "bf16x3f" (BF16 MFMA, forward
        # operands split hi/lo, the default: |dE| < 1e-4 Ha per molecule),
        # "bf16", "bf16x2", "fp16", "fp16x2", "bf16x3" or "fp32"
        self["MLPPrecision"] = "bf16x3f"


PARAMS = _Params()


def an1_grids(p=PARAMS):
    """Radial shells, angular shells and angles of SetANI1Param
    (TFMolInstanceDirect.py:3128-3164)."""
    rr, ra = p["AN1_r_Rc"], p["AN1_a_Rc"]
    nr, na, nt = p["AN1_num_r_Rs"], p["AN1_num_a_Rs"], p["AN1_num_a_As"]
    rs_r = np.array([rr * i / nr for i in range(nr)])
    rs_a = np.array([ra * i / na for i in range(na)])
    th = np.array([2.0 * math.pi * i / nt for i in range(nt)])
    return rs_r, rs_a, th
