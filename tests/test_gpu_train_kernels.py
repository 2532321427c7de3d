"""The training step's MLP kernels through the C ABI: the BF16-operand MFMA GEMM
against a plain PyTorch FP32 GEMM of the same BF16-rounded operands (every
transpose combination, sizes that are not tile multiples, each epilogue, split
K), the elementwise steps against their formulas, and the row gather /
scatter.  The products of BF16 operands are exact in FP32, so the GEMM differs
from the reference only by summation order: tolerance 1e-5 of sum |a||b|."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _lib():
    from tensormol_amd import _lib
    return _lib.lib(), _lib.ptr


def _r(t):
    return t.to(torch.bfloat16).float()


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(70, 130, 45), (64, 64, 32), (1, 513, 300), (257, 1, 96)])
def test_gemm_plain_and_bias(cuda, ta, tb, M, N, K):
    L, P = _lib()
    g = torch.Generator(device="cpu").manual_seed(M * 7 + N * 3 + K)
    A = torch.randn(M, K, generator=g).to(cuda)
    B = torch.randn(K, N, generator=g).to(cuda)
    bias = torch.randn(N, generator=g).to(cuda)
    As = A.t().contiguous() if ta else A          # stored layout
    Bs = B.t().contiguous() if tb else B
    lda = As.shape[1]
    ldb = Bs.shape[1]
    ref = _r(A) @ _r(B)
    scale = (_r(A).abs() @ _r(B).abs()).clamp_min(1e-6)
    for bz in (None, bias):
        C = torch.full((M, N + 3), float("nan"), device=cuda)
        st = torch.cuda.current_stream().cuda_stream
        L.tm_gemm_bf16(M, N, K, P(As), lda, ta, P(Bs), ldb, tb, P(C), N + 3, 0, P(bz), None,
                       None, 1.0, 1, None, st)
        torch.cuda.synchronize()
        r = ref + (0 if bz is None else bias)
        err = ((C[:, :N] - r).abs() / scale).max().item()
        assert err < 1e-5, (ta, tb, M, N, K, bz is not None, err)
        assert torch.isnan(C[:, N:]).all()   # nothing written past N


def test_gemm_split_k(cuda):
    L, P = _lib()
    g = torch.Generator(device="cpu").manual_seed(5)
    M, N, K = 513, 96, 4000
    A = torch.randn(K, M, generator=g).to(cuda)   # stored transposed (A(m, k) = A[k][m])
    B = torch.randn(K, N, generator=g).to(cuda)
    ref = _r(A).t() @ _r(B)
    scale = (_r(A).abs().t() @ _r(B).abs())
    outs = []
    for ns in (1, 3, 8):
        C = torch.zeros(M, N, device=cuda)
        work = torch.zeros(ns * M * N, device=cuda)
        L.tm_gemm_bf16(M, N, K, P(A), M, 1, P(B), N, 0, P(C), N, 0, None, None, None, 1.0, ns,
                       P(work), torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert ((C - ref).abs() / scale).max().item() < 1e-5
        outs.append(C)
    # the same split gives the same bits on a second call
    C2 = torch.zeros(M, N, device=cuda)
    work = torch.zeros(8 * M * N, device=cuda)
    L.tm_gemm_bf16(M, N, K, P(A), M, 1, P(B), N, 0, P(C2), N, 0, None, None, None, 1.0, 8,
                   P(work), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert torch.equal(C2, outs[-1])


def test_gemm_softplus_and_tangent_epilogues(cuda):
    L, P = _lib()
    g = torch.Generator(device="cpu").manual_seed(11)
    M, N, K, a = 100, 80, 72, 3.0
    A = torch.randn(M, K, generator=g).to(cuda) * 0.3
    B = torch.randn(K, N, generator=g).to(cuda) * 0.3
    bias = torch.randn(N, generator=g).to(cuda)
    s = torch.rand(M, N, generator=g).to(cuda)
    st = torch.cuda.current_stream().cuda_stream
    z = _r(A) @ _r(B) + bias
    C = torch.zeros(M, N, device=cuda)
    aux = torch.zeros(M, N, device=cuda)
    L.tm_gemm_bf16(M, N, K, P(A), K, 0, P(B), N, 0, P(C), N, 1, P(bias), None, P(aux), a, 1, None,
                   st)
    torch.cuda.synchronize()
    np.testing.assert_allclose(C.cpu(), (torch.nn.functional.softplus(a * z) / a).cpu(),
                               rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(aux.cpu(), torch.sigmoid(a * z).cpu(), rtol=1e-4, atol=1e-5)
    C.zero_(); aux.zero_()
    L.tm_gemm_bf16(M, N, K, P(A), K, 0, P(B), N, 0, P(C), N, 2, None, P(s), P(aux), 1.0, 1, None,
                   st)
    torch.cuda.synchronize()
    zt = _r(A) @ _r(B)
    np.testing.assert_allclose(C.cpu(), zt.cpu(), rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(aux.cpu(), (zt * s).cpu(), rtol=1e-4, atol=1e-5)


def test_train_elem_ops(cuda):
    L, P = _lib()
    g = torch.Generator(device="cpu").manual_seed(3)
    M, H, ld, a = 37, 50, 66, 2.5
    mk = lambda: torch.randn(M, ld, generator=g).to(cuda)
    s, x, zt, hd = torch.rand(M, ld, generator=g).to(cuda), mk(), mk(), mk()
    w4, yb = torch.randn(H, generator=g).to(cuda), torch.randn(M, generator=g).to(cuda)
    st = torch.cuda.current_stream().cuda_stream
    S, X, ZT, HD = s[:, :H], x[:, :H], zt[:, :H], hd[:, :H]
    ref = {0: S * w4, 1: S * X, 2: S * yb[:, None] * w4 + a * S * (1 - S) * ZT * w4,
           3: S * X + a * S * (1 - S) * ZT * HD}
    for op, r in ref.items():
        out = torch.zeros(M, ld, device=cuda)
        L.tm_train_elem(op, M, H, ld, a, P(s), P(x), P(zt), P(hd), P(w4), P(yb), P(out), st)
        torch.cuda.synchronize()
        np.testing.assert_allclose(out[:, :H].cpu(), r.cpu(), rtol=1e-6, atol=1e-6)
        assert (out[:, H:] == 0).all()


def test_rows_gather_scatter(cuda):
    L, P = _lib()
    idx = torch.tensor([5, 0, 3, 9], dtype=torch.int32, device=cuda)
    src = torch.arange(10 * 7, dtype=torch.float32, device=cuda).reshape(10, 7)
    out = torch.zeros(4, 9, device=cuda)
    st = torch.cuda.current_stream().cuda_stream
    L.tm_rows(0, 4, 7, P(idx), P(src), 7, P(out), 9, st)
    back = torch.zeros(10, 7, device=cuda)
    L.tm_rows(1, 4, 7, P(idx), P(out), 9, P(back), 7, st)
    torch.cuda.synchronize()
    assert torch.equal(out[:, :7], src[idx.long()])
    assert torch.equal(back[idx.long()], src[idx.long()])


@pytest.mark.parametrize("ta,tb,epi,ns", [(0, 0, 1, 1), (0, 1, 0, 1), (1, 0, 0, 4), (0, 0, 2, 1)])
def test_gemm_grouped_equals_single_calls(cuda, ta, tb, epi, ns):
    """17 GEMMs of mixed shapes (one empty; more than one launch's worth) in
    one grouped call give the bits of 17 single calls, and match the FP32
    GEMM of the rounded operands."""
    from tensormol_amd import _lib as TL
    L, P = _lib()
    g = torch.Generator(device="cpu").manual_seed(17 + 4 * ta + 2 * tb + epi)
    st = torch.cuda.current_stream().cuda_stream
    shapes = [(300, 256, 200), (129, 1, 70), (0, 64, 64), (40, 513, 130), (700, 128, 1000)]
    shapes = [shapes[i % len(shapes)] for i in range(17)]
    a = 2.0
    cases, groups, singles = [], [], []
    for (M, N, K) in shapes:
        A = torch.randn(M, K, generator=g).to(cuda) * 0.2
        B = torch.randn(K, N, generator=g).to(cuda) * 0.2
        As = A.t().contiguous() if ta else A
        Bs = B.t().contiguous() if tb else B
        lda, ldb = (max(M, 1) if ta else K), (K if tb else N)
        bias = torch.randn(N, generator=g).to(cuda) if epi == 1 else None
        s = torch.rand(M, N, generator=g).to(cuda) if epi == 2 else None
        outs = []
        for _ in range(2):
            C = torch.zeros(M, N, device=cuda)
            aux = torch.zeros(M, N, device=cuda) if epi else None
            work = torch.zeros(ns * M * N + 1, device=cuda) if ns > 1 else None
            outs.append((C, aux, work))
        cases.append((M, N, K, A, B, As, Bs, lda, ldb, bias, s, outs))
        ad = lambda t: None if t is None else t.data_ptr()
        C, aux, work = outs[0]
        groups.append(TL.GemmGroup(M, N, K, ad(As), lda, ad(Bs), ldb, ad(C), N, ad(bias), ad(s),
                                   ad(aux), ad(work)))
        C, aux, work = outs[1]
        singles.append((M, N, K, P(As), lda, ta, P(Bs), ldb, tb, P(C), N, epi, P(bias), P(s),
                        P(aux), a, ns, P(work), st))
    arr = (TL.GemmGroup * len(groups))(*groups)
    import ctypes
    L.tm_gemm_bf16_grouped(len(groups), ctypes.cast(arr, ctypes.c_void_p), ta, tb, epi, 0, a, ns,
                           st)
    for args in singles:
        L.tm_gemm_bf16(*args)
    torch.cuda.synchronize()
    for (M, N, K, A, B, As, Bs, lda, ldb, bias, s, outs) in cases:
        (C, aux, _), (C1, aux1, _) = outs
        assert torch.equal(C, C1), (M, N, K)
        if aux is not None:
            assert torch.equal(aux, aux1), (M, N, K)
        if M == 0:
            continue
        z = _r(A) @ _r(B)
        scale = (_r(A).abs() @ _r(B).abs()).clamp_min(1e-6)
        if epi == 1:
            z = z + bias
            np.testing.assert_allclose(C.cpu(), (torch.nn.functional.softplus(a * z) / a).cpu(),
                                       rtol=1e-4, atol=1e-5)
            np.testing.assert_allclose(aux.cpu(), torch.sigmoid(a * z).cpu(), rtol=1e-4, atol=1e-5)
        else:
            assert ((C - z).abs() / scale).max().item() < 1e-5, (M, N, K)
            if epi == 2:
                np.testing.assert_allclose(aux.cpu(), (C * s).cpu(), rtol=1e-6, atol=1e-7)


def test_gemm_fused_elementwise_epilogues(cuda):
    """The epilogues that carry the elementwise steps of the double backward
    against tm_train_elem's formulas on the FP32 GEMM of the rounded
    operands: softplus + out2 = sigmoid * w4[n] (op 0), tangent + out2 (op 2,
    C not written), and the reverse epilogue (op 3)."""
    import ctypes
    from tensormol_amd import _lib as TL
    L, P = _lib()
    g = torch.Generator(device="cpu").manual_seed(23)
    M, N, K, a, ld = 150, 140, 96, 2.5, 144
    st = torch.cuda.current_stream().cuda_stream
    A = torch.randn(M, K, generator=g).to(cuda) * 0.3
    B = torch.randn(K, N, generator=g).to(cuda) * 0.3
    bias = torch.randn(N, generator=g).to(cuda)
    s = torch.rand(M, ld, generator=g).to(cuda)
    w4 = torch.randn(N, generator=g).to(cuda)
    yb = torch.randn(M, generator=g).to(cuda)
    zt, hd = torch.randn(M, ld, generator=g).to(cuda), torch.randn(M, ld, generator=g).to(cuda)
    acc = _r(A) @ _r(B)
    S = s[:, :N]
    ad = lambda t: t.data_ptr()

    def run(epi, **f):
        grp = TL.GemmGroup(M=M, N=N, K=K, A=ad(A), lda=K, B=ad(B), ldb=N, ldc=ld, **f)
        arr = (TL.GemmGroup * 1)(grp)
        L.tm_gemm_bf16_grouped(1, ctypes.cast(arr, ctypes.c_void_p), 0, 0, epi, 0, a, 1, st)
        torch.cuda.synchronize()

    C, aux, out2 = (torch.full((M, ld), 7.0, device=cuda) for _ in range(3))
    run(1, C=ad(C), bias=ad(bias), aux=ad(aux), x1=ad(w4), out2=ad(out2))
    sg = torch.sigmoid(a * (acc + bias))
    np.testing.assert_allclose(aux[:, :N].cpu(), sg.cpu(), rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(out2[:, :N].cpu(), (sg * w4).cpu(), rtol=1e-4, atol=1e-5)
    assert (out2[:, N:] == 7.0).all()
    C, aux, out2 = (torch.full((M, ld), 7.0, device=cuda) for _ in range(3))
    run(2, C=None, s=ad(s), aux=ad(aux), x1=ad(w4), x2=ad(yb), out2=ad(out2))
    ref2 = S * yb[:, None] * w4 + a * S * (1 - S) * acc * w4
    np.testing.assert_allclose(aux[:, :N].cpu(), (acc * S).cpu(), rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(out2[:, :N].cpu(), ref2.cpu(), rtol=1e-4, atol=1e-5)
    assert (C == 7.0).all()
    C = torch.full((M, ld), 7.0, device=cuda)
    run(3, C=ad(C), s=ad(s), x1=ad(zt), x2=ad(hd))
    ref3 = S * acc + a * S * (1 - S) * zt[:, :N] * hd[:, :N]
    np.testing.assert_allclose(C[:, :N].cpu(), ref3.cpu(), rtol=1e-4, atol=1e-5)
    assert (C[:, N:] == 7.0).all()


def test_rows_grouped_equals_single_calls(cuda):
    import ctypes
    from tensormol_amd import _lib as TL
    L, P = _lib()
    st = torch.cuda.current_stream().cuda_stream
    src = torch.arange(40 * 9, dtype=torch.float32, device=cuda).reshape(40, 9)
    sets = []
    outs = []
    for k, idx in enumerate(([5, 0, 3, 9], [1, 2], [], [39, 20, 7])):
        it = torch.tensor(idx, dtype=torch.int32, device=cuda)
        o = torch.zeros(max(len(idx), 1), 11, device=cuda)
        outs.append((it, o))
        sets.append(TL.RowsGroup(len(idx), 9, it.data_ptr(), src.data_ptr(), 9, o.data_ptr(), 11))
    arr = (TL.RowsGroup * len(sets))(*sets)
    L.tm_rows_grouped(len(sets), ctypes.cast(arr, ctypes.c_void_p), 0, st)
    torch.cuda.synchronize()
    for it, o in outs:
        if it.numel():
            assert torch.equal(o[:it.numel(), :9], src[it.long()])
            assert (o[:, 9:] == 0).all()


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K,pad", [(200, 136, 200, 8), (70, 130, 45, 3)])
def test_gemm_bf16_storage_equals_fp32_staging(cuda, ta, tb, M, N, K, pad):
    """Operands stored as BF16 (io bit 0) give the bits of the same operands in
    FP32, which the kernel rounds to BF16 when it stages them (16-byte loads
    with the padded strides, element loads with the odd ones); outputs written
    as BF16 (io bits 1-3) are the FP32 outputs rounded to nearest even."""
    import ctypes
    from tensormol_amd import _lib as TL
    L, P = _lib()
    g = torch.Generator(device="cpu").manual_seed(100 + M + 7 * ta + 3 * tb)
    st = torch.cuda.current_stream().cuda_stream
    A = torch.randn(M, K, generator=g).to(cuda) * 0.3
    B = torch.randn(K, N, generator=g).to(cuda) * 0.3
    bias = torch.randn(N, generator=g).to(cuda)
    w4 = torch.randn(N, generator=g).to(cuda)
    As = A.t() if ta else A
    Bs = B.t() if tb else B
    lda, ldb = As.shape[1] + pad, Bs.shape[1] + pad
    A32 = torch.zeros(As.shape[0], lda, device=cuda); A32[:, :As.shape[1]] = As
    B32 = torch.zeros(Bs.shape[0], ldb, device=cuda); B32[:, :Bs.shape[1]] = Bs
    A16, B16 = A32.bfloat16(), B32.bfloat16()
    ldc = N + pad

    def run(Ap, Bp, io, epi):
        outs = [torch.zeros(M, ldc, device=cuda, dtype=torch.bfloat16 if io & bit else torch.float32)
                for bit in (2, 4, 8)]
        f = dict(M=M, N=N, K=K, A=Ap.data_ptr(), lda=lda, B=Bp.data_ptr(), ldb=ldb,
                 C=outs[0].data_ptr(), ldc=ldc)
        if epi == 1:
            f.update(bias=bias.data_ptr(), aux=outs[1].data_ptr(), x1=w4.data_ptr(),
                     out2=outs[2].data_ptr())
        arr = (TL.GemmGroup * 1)(TL.GemmGroup(**f))
        L.tm_gemm_bf16_grouped(1, ctypes.cast(arr, ctypes.c_void_p), ta, tb, epi, io, 1.5, 1, st)
        torch.cuda.synchronize()
        return outs

    for epi in (0, 1):
        ref = run(A32, B32, 0, epi)
        same = run(A16, B16, 1, epi)
        for r, o in zip(ref, same):
            assert torch.equal(r, o)
        half = run(A16, B16, 1 | 2 | 4 | 8, epi)
        for r, o in zip(ref, half):
            assert torch.equal(r.bfloat16(), o)


def test_rows_bf16_and_cast(cuda):
    import ctypes
    from tensormol_amd import _lib as TL
    L, P = _lib()
    st = torch.cuda.current_stream().cuda_stream
    g = torch.Generator(device="cpu").manual_seed(9)
    src = torch.randn(40, 9, generator=g).to(cuda)
    it = torch.tensor([5, 0, 3, 39], dtype=torch.int32, device=cuda)
    out = torch.zeros(4, 12, dtype=torch.bfloat16, device=cuda)
    arr = (TL.RowsGroup * 1)(TL.RowsGroup(4, 9, it.data_ptr(), src.data_ptr(), 9, out.data_ptr(), 12, 1))
    L.tm_rows_grouped(1, ctypes.cast(arr, ctypes.c_void_p), 0, st)
    xs = [torch.randn(n, generator=g).to(cuda) * 10 for n in (1, 1000, 70000)]
    ys = [torch.zeros(x.numel(), dtype=torch.bfloat16, device=cuda) for x in xs]
    carr = (TL.CastGroup * 3)(*[TL.CastGroup(x.data_ptr(), y.data_ptr(), x.numel()) for x, y in zip(xs, ys)])
    L.tm_cast_bf16_grouped(3, ctypes.cast(carr, ctypes.c_void_p), st)
    torch.cuda.synchronize()
    assert torch.equal(out[:, :9], src[it.long()].bfloat16())
    assert (out[:, 9:] == 0).all()
    for x, y in zip(xs, ys):
        assert torch.equal(y, x.bfloat16())


def test_fm_loss_against_torch_formulas(cuda):
    """tm_fm_loss (molecule energies, loss, ybar, dLoss/dgrad) against the same
    formulas in FP64 torch on ragged molecules, one of them empty."""
    L, P = _lib()
    g = torch.Generator(device="cpu").manual_seed(31)
    sizes = [3, 70, 1, 0, 130, 64]
    ms = torch.tensor([0] + list(np.cumsum(sizes)), dtype=torch.int32)
    N, B = int(ms[-1]), len(sizes)
    mol = torch.repeat_interleave(torch.arange(B), torch.tensor(sizes))
    eps = torch.randn(N, 1, generator=g)
    E_ref, w_e, w_f = torch.randn(B, generator=g, dtype=torch.float64), \
        torch.rand(B, generator=g, dtype=torch.float64), torch.rand(B, generator=g, dtype=torch.float64)
    grad, F_ref = torch.randn(N, 3, generator=g, dtype=torch.float64), torch.randn(N, 3, generator=g, dtype=torch.float64)
    E_t = torch.zeros(B, dtype=torch.float64).index_add(0, mol, eps[:, 0].double())
    F = -grad
    wf = w_f[mol][:, None]
    L_t = (w_e * (E_t - E_ref) ** 2).sum() + (wf * (F - F_ref) ** 2).sum()
    ybar_t = (2.0 * w_e * (E_t - E_ref))[mol].float().reshape(N, 1)
    gg_t = -2.0 * wf * (F - F_ref)
    d = lambda t: t.to(cuda).contiguous()
    E, part, Lo = (torch.empty(n, dtype=torch.float64, device=cuda) for n in (B, B, 1))
    ybar = torch.empty(N, 1, device=cuda)
    gg = torch.empty(N, 3, dtype=torch.float64, device=cuda)
    ms_d, eps_d, Er, we, gr, Fr, wfd = map(d, (ms, eps, E_ref, w_e, grad, F_ref, w_f))
    L.tm_fm_loss(B, P(ms_d), P(eps_d), P(Er), P(we), P(gr), P(Fr), P(wfd), P(E), P(ybar), P(gg),
                 P(part), P(Lo), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    np.testing.assert_allclose(E.cpu().numpy(), E_t.numpy(), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(Lo.cpu().numpy()[0], L_t.item(), rtol=1e-12)
    np.testing.assert_allclose(ybar.cpu().numpy(), ybar_t.numpy(), rtol=1e-6)
    np.testing.assert_allclose(gg.cpu().numpy(), gg_t.numpy(), rtol=1e-12, atol=1e-15)


def test_gemm_row_map_scatters_rows(cuda):
    """crow: row m of the product lands in row crow[m] of C (plain epilogue,
    with bias), the rows not in the map untouched; equal to the unmapped
    result scattered."""
    import ctypes
    from tensormol_amd import _lib as TL
    L, P = _lib()
    g = torch.Generator(device="cpu").manual_seed(41)
    M, N, K = 150, 70, 96
    A = torch.randn(M, K, generator=g).to(cuda)
    B = torch.randn(K, N, generator=g).to(cuda)
    bias = torch.randn(N, generator=g).to(cuda)
    rows = torch.randperm(400, generator=g)[:M].to(torch.int32).to(cuda)
    st = torch.cuda.current_stream().cuda_stream
    C0 = torch.zeros(M, N, device=cuda)
    C1 = torch.full((400, N), 3.0, device=cuda)
    for C, crow in ((C0, None), (C1, rows)):
        f = dict(M=M, N=N, K=K, A=A.data_ptr(), lda=K, B=B.data_ptr(), ldb=N, C=C.data_ptr(),
                 ldc=N, bias=bias.data_ptr(), crow=None if crow is None else crow.data_ptr())
        arr = (TL.GemmGroup * 1)(TL.GemmGroup(**f))
        L.tm_gemm_bf16_grouped(1, ctypes.cast(arr, ctypes.c_void_p), 0, 0, 0, 0, 1.0, 1, st)
    torch.cuda.synchronize()
    assert torch.equal(C1[rows.long()], C0)
    mask = torch.ones(400, dtype=torch.bool, device=cuda)
    mask[rows.long()] = False
    assert (C1[mask] == 3.0).all()
