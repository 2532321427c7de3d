"""MI355X-native BP+EE (TensorMol 0.1) energy/force engine.

The drop-in names of the reference's BP+EE API are exported here; the
numerical work runs in libtmamd.so (HIP, gfx950).
"""
from .params import PARAMS, BOHRPERA, JOULEPERHARTREE, KJPERHARTREE, IDEALGASR  # noqa: F401
from .containers import Mol, MSet  # noqa: F401
from .model import BPEEModel  # noqa: F401

_LAZY = {
    "TFMolManage": "manage", "MolDigester": "manage", "TensorMolData_BP_Direct_EE_WithEle": "manage",
    "BPEEInstance": "manage",
    "NeighborListSet": "neighbors", "NeighborListSetWithImages": "neighbors",
    "NeighborList": "neighbors", "Make_NListNaive": "neighbors",
    "VelocityVerlet": "md", "VelocityVerletStep": "md", "Thermostat": "md",
    "NoseThermostat": "md", "KineticEnergy": "md", "DeviceMD": "md",
    "GeomOptimizer": "opt", "ConjGradient": "opt", "RemoveInvariantForce": "opt",
    "DeviceFIRE": "opt",
    "BPEEEngine": "engine",
    "Server": "serving",
}


def __getattr__(name):
    if name in _LAZY:
        import importlib
        mod = importlib.import_module(f".{_LAZY[name]}", __name__)
        return getattr(mod, name)
    raise AttributeError(name)
