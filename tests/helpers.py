"""Shared builders for the parity tests: seeded systems, models and oracle runs."""
import functools

import numpy as np

from oracle import oracle as O
from tensormol_amd import workloads as W
from tensormol_amd.model import BPEEModel


def water_model(seed=0):
    return BPEEModel([1, 8], dsf_alpha=0.2, seed=seed)


def chno_model(seed=0, hidden=(512, 512, 512)):
    return BPEEModel([1, 6, 7, 8], dsf_alpha=0.2, seed=seed, hidden=list(hidden))


@functools.lru_cache(maxsize=None)
def system(name):
    """Returns (Z, xyz, mol_start, boxes)."""
    if name == "water64":
        Z, x = W.water_cluster(64, seed=0)
        return Z, x, np.array([0, len(Z)]), None
    if name == "waterbox10":
        Z, x, box = W.water_box(nside=10, edge=31.06, seed=5)
        return Z, x, np.array([0, len(Z)]), box[None, :]
    if name == "chains":
        mols = W.chain_set(W.request_sizes(12, seed=7), seed=7)
        Z, x, ms = W.pack(mols)
        return Z, x, ms, None
    if name == "chno_small":
        Z, x = W.chno_sphere(natom=1200, seed=11)
        return Z, x, np.array([0, len(Z)]), None
    if name in GOLDEN:
        mols = golden_mols(GOLDEN[name][0])[: GOLDEN[name][1]]
        return W.pack([(m.atoms.astype(np.int32), m.coords) for m in mols])[:3] + (None,)
    raise KeyError(name)


# geometries from the reference's datasets/ (tests/golden): file, molecules used
GOLDEN = {"morphine": ("morphine", 1), "water10": ("water10", 1), "peptide188": ("2mzx_tm_opt", 1),
          "johnsonmols": ("johnsonmols", 10)}


def golden_mols(fname):
    import os
    from tensormol_amd.containers import MSet
    ms = MSet(fname, path_=os.path.join(os.path.dirname(__file__), "golden"), center_=False)
    ms.ReadXYZ()
    return ms.mols


@functools.lru_cache(maxsize=None)
def oracle_eval(name, modelname, need_forces=True):
    Z, x, ms, boxes = system(name)
    model = water_model() if modelname == "water" else chno_model()
    sp = O.SymParams(model.eles)
    r = O.evaluate(x, Z, ms, sp, model.energy_nets, model.charge_nets, alpha=model.alpha,
                   dsf_alpha=model.dsf_alpha, has_vdw=model.has_vdw, boxes=boxes,
                   need_forces=need_forces)
    return r


def model_for(modelname):
    return water_model() if modelname == "water" else chno_model()
