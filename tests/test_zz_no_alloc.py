"""tm_mol_sums_static, the entry point that allocates nothing, against
tm_mol_sums with a caller workspace: bit-identical per-molecule sums on
consecutive calls (the second smaller, so the static array's leftover
partials from the first call must not leak in), on the many-block path of one
large system and on the one-block-per-molecule path of a batch, and a clear
error above its size limit."""
import ctypes

import numpy as np
import pytest
import torch

from tensormol_amd import _lib

pytestmark = pytest.mark.gpu

MAX_ATOMS = 1 << 22


def _sources(N, B, seed, dev):
    """Three sources of 1 + 2 + 3 columns (float32, float64 with ld 3, float32
    with ld 4) over N atoms split into B molecules."""
    rng = np.random.default_rng(seed)
    cuts = np.sort(rng.choice(np.arange(1, N), size=B - 1, replace=False)) if B > 1 else []
    ms = np.concatenate([[0], cuts, [N]]).astype(np.int32)
    a = torch.tensor(rng.normal(size=(N, 1)), dtype=torch.float32, device=dev)
    b = torch.tensor(rng.normal(size=(N, 3)) * 1e3, dtype=torch.float64, device=dev)
    c = torch.tensor(rng.normal(size=(N, 4)), dtype=torch.float32, device=dev)
    srcs = ((a, 0, 1, 1), (b, 1, 2, 3), (c, 0, 3, 4))
    return torch.tensor(ms, device=dev), srcs


def _run(fn, B, ms, N, srcs, work, dev):
    outs = [torch.full((B, k), np.nan, dtype=torch.float64, device=dev) for _, _, k, _ in srcs]
    n = len(srcs)
    ins = (ctypes.c_void_p * n)(*[t.data_ptr() for t, *_ in srcs])
    f64 = (ctypes.c_int * n)(*[f for _, f, _, _ in srcs])
    cols = (ctypes.c_int * n)(*[k for _, _, k, _ in srcs])
    ld = (ctypes.c_int * n)(*[l for *_, l in srcs])
    op = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs])
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if work is None:
        assert fn(B, _lib.ptr(ms), N, n, ins, f64, cols, ld, op, st) == 0
    else:
        assert fn(B, _lib.ptr(ms), N, n, ins, f64, cols, ld, op, _lib.ptr(work), st) == 0
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in outs]


@pytest.mark.parametrize("sizes", [((300_000, 1), (20_000, 1)), ((50_000, 37), (3_000, 5))])
def test_static_sums_equal_workspace_sums(cuda, sizes):
    L = _lib.lib()
    for k, (N, B) in enumerate(sizes):
        print(f"static sums: call {k}, N {N}, B {B}", flush=True)
        ms, srcs = _sources(N, B, 10 + k, cuda)
        work = torch.zeros(1 + ((N + 1023) // 1024) * 6, dtype=torch.float64, device=cuda)
        ref = _run(L.tm_mol_sums, B, ms, N, srcs, work, cuda)
        got = _run(L.tm_mol_sums_static, B, ms, N, srcs, None, cuda)
        for r, g in zip(ref, got):
            assert np.isfinite(g).all()
            np.testing.assert_array_equal(g, r)
        # and against a host FP64 sum of the same values
        msn = ms.cpu().numpy()
        t0 = srcs[1][0].cpu().numpy()
        host = np.array([t0[msn[m]:msn[m + 1], :2].sum(0) for m in range(B)])
        np.testing.assert_allclose(got[1], host, rtol=1e-12, atol=1e-6)


def test_static_sums_rejects_above_limit(cuda):
    L = _lib.lib()
    with pytest.raises(RuntimeError, match="static scratch"):
        L.tm_mol_sums_static(1, None, MAX_ATOMS + 1, 1, None, None, None, None, None, None)
