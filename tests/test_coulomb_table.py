"""The pair-function tables of the tabulated Coulomb sweep (tm_coulomb_table, a
host routine of the C-ABI library, so these run without a GPU) against the
oracle's float64 DSF-elu and vdW kernels (RawSymFunc.py:1307-1414).

Each table is evaluated here in float64 from its float32 cubic pieces, as
the sweep does in float32: values within 2e-8 of the oracle's (Ha per unit
charge product for K, Ha for V; |K| <= 0.03, |V| <= 1.2e-4 over the cutoff),
r-derivatives against central differences of the oracle within 2e-5 of the
largest derivative (K) and 2e-3 (V, whose switch has a kink in its second
derivative between nodes).  The tables' element-pair order and their zeroing by
AddEcc / HasVdw are checked too.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tensormol_amd import _lib
from tensormol_amd.engine import dsf_constants
from tensormol_amd.model import BPEEModel
from tensormol_amd.params import BOHRPERA

ELES = [1, 6, 7, 8]
NINT = 256          # the interval count the engine uses
RC = 15.0


@pytest.fixture(scope="module")
def lib():
    return _lib.lib()


def _tables(lib, model, add_ecc=1, has_vdw=1, nint=NINT):
    nele = len(model.eles)
    ntab = 1 + nele * (nele + 1) // 2
    out = np.zeros(4 * ntab * nint, dtype=np.float32)
    scale = np.zeros(2)
    c6, r0 = model.vdw_tables()
    consts = dsf_constants(model.dsf_alpha, model.elu_width, model.ee_off)
    lib.tm_coulomb_table(consts.ctypes.data_as(ctypes.c_void_p), model.ee_off,
                         model.poly_width * BOHRPERA, add_ecc, has_vdw, nele,
                         np.ascontiguousarray(c6).ctypes.data_as(ctypes.c_void_p),
                         np.ascontiguousarray(r0).ctypes.data_as(ctypes.c_void_p), nint,
                         out.ctypes.data_as(ctypes.c_void_p), scale.ctypes.data_as(ctypes.c_void_p))
    return out.reshape(ntab, nint, 4).astype(np.float64), scale


def _eval(tabs, rA, t_index=0):
    tab, (inv_h, u0) = tabs[0][t_index], tabs[1]
    u = rA * inv_h + u0
    k = np.minimum(u.astype(np.int64), len(tab) - 1)
    t = u - k
    c = tab[k]
    v = c[:, 0] + t * (c[:, 1] + t * (c[:, 2] + t * c[:, 3]))
    dv = (c[:, 1] + t * (2 * c[:, 2] + 3 * t * c[:, 3])) * inv_h
    return v, dv


def _k_ref(model, rA):
    c = O.dsf_consts(model.dsf_alpha, model.elu_width, model.ee_off)
    return O.dsf_kernel(torch.tensor(rA * BOHRPERA, dtype=torch.float64), c).numpy()


def _v_ref(zi, zj, rA):
    c6, r0 = O.vdw_params([zi, zj])
    t = lambda v: torch.tensor(v, dtype=torch.float64)
    return O.vdw_kernel(t(rA * BOHRPERA), t(c6[0]), t(c6[1]), t(r0[0]), t(r0[1])).numpy()


def _pair_tables(nele):
    idx, t = {}, 1
    for a in range(nele):
        for b in range(a, nele):
            idx[(a, b)] = idx[(b, a)] = t
            t += 1
    return idx


def test_tables_match_oracle_kernels(lib):
    model = BPEEModel(ELES, dsf_alpha=0.2, seed=0)
    tab = _tables(lib, model)
    rA = np.random.default_rng(0).uniform(0.3, RC * (1 - 1e-9), 20000)
    h = 1e-6
    v, dv = _eval(tab, rA)
    ref = _k_ref(model, rA)
    dref = (_k_ref(model, rA + h) - _k_ref(model, rA - h)) / (2 * h)
    print(f"K: max |dK| {np.abs(v - ref).max():.2e}, max |d(dK/dr)| {np.abs(dv - dref).max():.2e} "
          f"(max |dK/dr| {np.abs(dref).max():.2e})")
    assert np.abs(v - ref).max() < 2e-8
    assert np.abs(dv - dref).max() < 2e-5 * np.abs(dref).max()
    idx = _pair_tables(len(ELES))
    for a, za in enumerate(ELES):
        for b, zb in enumerate(ELES):
            v, dv = _eval(tab, rA, idx[(a, b)])
            ref = _v_ref(za, zb, rA)
            dref = (_v_ref(za, zb, rA + h) - _v_ref(za, zb, rA - h)) / (2 * h)
            err, derr = np.abs(v - ref).max(), np.abs(dv - dref).max()
            assert err < 2e-8, (za, zb, err)
            assert derr < 2e-3 * np.abs(dref).max(), (za, zb, derr, np.abs(dref).max())


def test_tables_vanish_at_cutoff_and_switch_off(lib):
    model = BPEEModel(ELES, dsf_alpha=0.2, seed=0)
    tabs = _tables(lib, model)
    tab, (inv_h, u0) = tabs
    # the grid has a node on the elu matching radius, and the shifted-force
    # kernel and its slope vanish at the cutoff
    m = (model.elu_width * inv_h + u0)
    assert abs(m - round(m)) < 1e-9
    v, dv = _eval(tabs, np.array([RC, RC - 1e-7]))
    assert abs(v[0]) < 1e-9 and abs(dv[0]) < 1e-8
    # V = 0 at r <= 0 (r^6 in the numerator and the switch S(0) = 0)
    assert np.all(tab[1:, 0, 0] == 0.0)
    assert 0.0 <= u0 < 1.0 and (RC * inv_h + u0) < tab.shape[1]
    off = _tables(lib, model, add_ecc=0)[0]
    assert np.all(off[0] == 0.0) and np.array_equal(off[1:], tab[1:])
    novdw = _tables(lib, model, has_vdw=0)[0]
    assert np.all(novdw[1:] == 0.0) and np.array_equal(novdw[0], tab[0])


def test_table_argument_errors(lib):
    model = BPEEModel([1, 8], dsf_alpha=0.2, seed=0)
    with pytest.raises(RuntimeError, match="nint"):
        _tables(lib, model, nint=8)
