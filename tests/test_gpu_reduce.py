"""Per-molecule sums through the C ABI (tm_mol_sum_f32 / tm_mol_sum_f64).

The reductions are restatements of reduce_sum over a molecule's atoms
(RawSymFunc.py:1354-1359, TFMolInstanceDirect.py:5212).  Checked against
exactly rounded sums (math.fsum) of the same inputs: one large system summed
by many blocks in one launch (completion counter in the scratch buffer, so
back-to-back calls, a smaller call after a larger one, and strided inputs are
covered), one system without scratch (single block), and many molecules (one
block each).  Results must also be bitwise repeatable.
"""
import math

import numpy as np
import pytest
import torch

from tensormol_amd import _lib

pytestmark = pytest.mark.gpu


def _ref(x, starts, mean):
    out = []
    for m in range(len(starts) - 1):
        lo, hi = starts[m], starts[m + 1]
        row = [math.fsum(x[lo:hi, c].astype(np.float64)) for c in range(x.shape[1])]
        if mean:
            row = [v / (hi - lo) if hi > lo else 0.0 for v in row]
        out.append(row)
    return np.array(out, dtype=np.float64)


def _call(L, dtype, starts, x, C, mean, scratch, dev):
    B = len(starts) - 1
    st = torch.tensor(starts, dtype=torch.int32, device=dev)
    xin = torch.tensor(x, dtype=dtype, device=dev)
    out = torch.full((B, C), float("nan"), dtype=torch.float64, device=dev)
    fn = L.tm_mol_sum_f32 if dtype == torch.float32 else L.tm_mol_sum_f64
    fn(B, _lib.ptr(st), C, _lib.ptr(xin), x.shape[1], _lib.ptr(out), int(mean), _lib.ptr(scratch),
       int(starts[-1]), None)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_single_system_many_blocks(cuda, dtype):
    L = _lib.lib()
    rng = np.random.default_rng(7)
    scratch = torch.zeros(1 + 300 * 8, dtype=torch.float64, device=cuda)
    # N, columns used, stored columns (ldin), mean; the second call is smaller
    cases = [(12288, 3, 3, 0), (2049, 1, 2, 1), (300000, 5, 7, 0), (12288, 2, 2, 1),
             (4097, 4, 4, 0)]
    for N, C, ld, mean in cases:
        x = rng.standard_normal((N, ld)) * 10.0 + 0.5
        if dtype == torch.float32:
            x = x.astype(np.float32)
        got = _call(L, dtype, [0, N], x, C, mean, scratch, cuda)
        want = _ref(x[:, :C], [0, N], mean)
        scale = np.abs(x[:, :C]).astype(np.float64).sum(axis=0) / (N if mean else 1)
        assert np.all(np.abs(got - want) <= 1e-13 * scale + 1e-300), (N, C, got, want)
        again = _call(L, dtype, [0, N], x, C, mean, scratch, cuda)
        assert np.array_equal(got, again), "not bitwise repeatable"
    # the completion counter is back at zero after every call
    assert scratch[:1].view(torch.int32)[0].item() == 0


def test_single_system_without_scratch(cuda):
    L = _lib.lib()
    rng = np.random.default_rng(3)
    x = rng.standard_normal((5000, 2))
    got = _call(L, torch.float64, [0, 5000], x, 2, 0, None, cuda)
    want = _ref(x, [0, 5000], 0)
    assert np.allclose(got, want, rtol=0, atol=1e-11)


def test_many_molecules(cuda):
    L = _lib.lib()
    rng = np.random.default_rng(5)
    sizes = rng.integers(0, 300, size=37)
    starts = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    x = rng.standard_normal((int(starts[-1]), 3))
    scratch = torch.zeros(1 + 8 * 64, dtype=torch.float64, device=cuda)
    for mean in (0, 1):
        got = _call(L, torch.float64, list(starts), x, 3, mean, scratch, cuda)
        want = _ref(x, list(starts), mean)
        assert np.allclose(got, want, rtol=0, atol=1e-11)


@pytest.mark.parametrize("sizes", [[12288], [2500], list(range(1, 40))])
def test_fused_sums_match_single_sums(cuda, sizes):
    """tm_mol_sums over (float32 [N], float64 [N,2], float64 [N,3]) gives the
    exactly rounded per-molecule sums, like three single-source calls."""
    import ctypes
    L = _lib.lib()
    rng = np.random.default_rng(11)
    starts = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    N, B = int(starts[-1]), len(sizes)
    a = rng.standard_normal((N, 1)).astype(np.float32)
    b = rng.standard_normal((N, 2)) * 3.0
    c = rng.standard_normal((N, 3)) - 0.25
    st = torch.tensor(starts, dtype=torch.int32, device=cuda)
    ta = torch.tensor(a, device=cuda)
    tb = torch.tensor(b, device=cuda)
    tc = torch.tensor(c, device=cuda)
    outs = [torch.full((B, k), float("nan"), dtype=torch.float64, device=cuda) for k in (1, 2, 3)]
    scratch = torch.zeros(1 + 6 * (N // 1024 + 1), dtype=torch.float64, device=cuda)
    ins = (ctypes.c_void_p * 3)(ta.data_ptr(), tb.data_ptr(), tc.data_ptr())
    ous = (ctypes.c_void_p * 3)(*[o.data_ptr() for o in outs])
    f64 = (ctypes.c_int * 3)(0, 1, 1)
    cols = (ctypes.c_int * 3)(1, 2, 3)
    ld = (ctypes.c_int * 3)(1, 2, 3)
    for _ in range(2):
        L.tm_mol_sums(B, _lib.ptr(st), N, 3, ins, f64, cols, ld, ous, _lib.ptr(scratch), None)
        torch.cuda.synchronize()
        for x, o in zip((a, b, c), outs):
            want = _ref(x, list(starts), 0)
            scale = np.abs(x).astype(np.float64).sum()
            assert np.all(np.abs(o.cpu().numpy() - want) <= 1e-13 * scale)
    assert scratch[:1].view(torch.int32)[0].item() == 0


@pytest.mark.parametrize("N,n", [(3000, 6), (12288, 21), (20000, 40)])
def test_cell_build_leaves_counts_zero(cuda, N, n):
    """tm_cell_build needs cell_count zero on entry and leaves it zero; cells
    list every atom once, in index order inside each cell, and cell_of is
    floor((x - origin) / edge) wrapped into the grid.  64,000 cells take the
    two-launch scan (four 16384-entry tiles), 216 and 9,261 the single block."""
    L = _lib.lib()
    rng = np.random.default_rng(4)
    edge = 20.0
    pos = torch.tensor(rng.uniform(0, edge, (N, 3)), dtype=torch.float64, device=cuda)
    mol = torch.zeros(N, dtype=torch.int32, device=cuda)
    i32 = dict(dtype=torch.int32, device=cuda)
    ncell = n ** 3
    cell_of, count = torch.zeros(N, **i32), torch.zeros(ncell, **i32)
    start, fill = torch.zeros(ncell + 1, **i32), torch.zeros(N, **i32)
    atoms, cpos = torch.zeros(N, **i32), torch.zeros(N, **i32)
    posc = torch.zeros(N, 3, dtype=torch.float64, device=cuda)
    flags = torch.zeros(8, **i32)
    cs = edge / n
    for _ in range(2):
        L.tm_cell_build(N, _lib.ptr(pos), _lib.ptr(mol), 0.0, 0.0, 0.0, cs, cs, cs, n, n, n, 1,
                        _lib.ptr(cell_of), _lib.ptr(count), _lib.ptr(start), _lib.ptr(fill),
                        _lib.ptr(atoms), _lib.ptr(cpos), _lib.ptr(posc), None, _lib.ptr(flags),
                        None)
        torch.cuda.synchronize()
        assert int(count.abs().sum().item()) == 0
        a, s, co = atoms.cpu().numpy(), start.cpu().numpy(), cell_of.cpu().numpy()
        q = np.mod(np.floor((pos.cpu().numpy() - 0.0) / cs).astype(np.int64), n)
        np.testing.assert_array_equal(co, q[:, 0] + n * (q[:, 1] + n * q[:, 2]))
        assert sorted(a.tolist()) == list(range(N))
        assert s[0] == 0 and s[-1] == N and np.all(np.diff(s) >= 0)
        np.testing.assert_array_equal(np.diff(s), np.bincount(co, minlength=ncell))
        for c in np.flatnonzero(np.diff(s)):
            seg = a[s[c]:s[c + 1]]
            assert np.all(np.diff(seg) > 0) and np.all(co[seg] == c)
        assert np.array_equal(cpos.cpu().numpy()[a], np.arange(N))
        assert torch.equal(posc, pos[atoms.long()])
