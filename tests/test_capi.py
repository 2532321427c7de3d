"""C-ABI library checks that need no GPU: it loads, exports every symbol the
header declares, and reports argument errors through tm_last_error."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        names |= set(re.findall(r"^(?:int|const char\*)\s+(tm_\w+)\s*\(", txt, re.M))
    return names


@pytest.fixture(scope="module")
def lib():
    from tensormol_amd import build, _lib
    if not os.path.exists(_lib.LIB_PATH):
        build.build()
    return _lib.lib()


def test_header_declares_entry_points():
    names = _declared()
    for n in ("tm_nl_build", "tm_sym_forward", "tm_sym_adjoint", "tm_mlp_mfma", "tm_coulomb",
              "tm_vv_kick", "tm_nh_kick", "tm_fire_update"):
        assert n in names


def test_every_declared_symbol_is_exported(lib):
    for n in _declared():
        assert hasattr(lib.dll, n), f"{n} declared in include/ but not exported"


def test_python_binding_covers_header(lib):
    from tensormol_amd._lib import SIGNATURES
    missing = {n for n in _declared() if n not in SIGNATURES and n != "tm_last_error"}
    assert not missing, missing
    for n in SIGNATURES:
        assert n in _declared(), f"{n} bound in _lib but not declared in the header"


def test_version_and_argument_errors(lib):
    assert lib.tm_version() == 1
    with pytest.raises(RuntimeError, match="bad nele"):
        lib.tm_nl_build(10, None, None, None, None, None, None, None, None, None, None, None, 1, 1, 1,
                        1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0, 0, 4.6, 1.0, 1.0, 128, 48, None, None, None,
                        None, None, None, None, None, None)
    with pytest.raises(RuntimeError, match="3 hidden layers"):
        out = (ctypes.c_long * 10)()
        lib.tm_mlp_mfma_pack_info(256, 512, 2, 0, out)
    out = (ctypes.c_long * 10)()
    lib.tm_mlp_mfma_pack_info(256, 512, 3, 0, out)
    assert out[8] == 1 and out[9] == 1 and out[0] > 0
    # split forward weights (mode 3): two forward planes, one backward plane
    out3 = (ctypes.c_long * 10)()
    lib.tm_mlp_mfma_pack_info(256, 512, 3, 3, out3)
    assert out3[8] == 2 and out3[9] == 1
    assert out3[3] - out3[2] == 2 * (out[3] - out[2])   # f1 holds two planes
    assert out3[6] - out3[5] == out[6] - out[5]         # b3 holds one


def test_binding_argument_counts_match_header():
    """Each ctypes signature has as many arguments as the C declaration."""
    from tensormol_amd._lib import SIGNATURES
    txt = "\n".join(open(h).read() for h in glob.glob(os.path.join(ROOT, "include", "*.h")))
    for name, args in SIGNATURES.items():
        m = re.search(r"(?:int|const char\*)\s+" + name + r"\s*\(([^;]*?)\);", txt, re.S)
        assert m, name
        params = m.group(1).strip()
        n = 0 if params in ("", "void") else params.count(",") + 1
        assert n == len(args), f"{name}: header has {n} parameters, binding {len(args)}"


def test_training_kernel_argument_errors(lib):
    """The training-step entry points reject bad arguments on the host, before
    anything is launched (so these run without a GPU)."""
    from tensormol_amd import _lib as TL
    fake = 4096   # a non-NULL address that is never dereferenced: every case fails validation

    def call(groups, epi=0, io=0, nsplit=1):
        arr = (TL.GemmGroup * len(groups))(*groups)
        lib.tm_gemm_bf16_grouped(len(groups), ctypes.cast(arr, ctypes.c_void_p), 0, 0, epi, io,
                                 1.0, nsplit, None)

    base = dict(M=8, N=8, K=8, A=fake, lda=8, B=fake, ldb=8, C=fake, ldc=8)
    with pytest.raises(RuntimeError, match="bad io flags"):
        call([TL.GemmGroup(**base)], io=16)
    with pytest.raises(RuntimeError, match="bad epilogue"):
        call([TL.GemmGroup(**base)], epi=4)
    with pytest.raises(RuntimeError, match="FP32 C"):
        call([TL.GemmGroup(**base, work=fake)], io=2, nsplit=2)
    with pytest.raises(RuntimeError, match="negative size"):
        call([TL.GemmGroup(**dict(base, M=-1))])
    with pytest.raises(RuntimeError, match="softplus epilogue needs bias and aux"):
        call([TL.GemmGroup(**base)], epi=1)
    with pytest.raises(RuntimeError, match="reverse epilogue"):
        call([TL.GemmGroup(**base, s=fake)], epi=3)
    with pytest.raises(RuntimeError, match="row map"):
        call([TL.GemmGroup(**base, bias=fake, aux=fake, crow=fake)], epi=1)
    with pytest.raises(RuntimeError, match="out2 needs"):
        call([TL.GemmGroup(**base, out2=fake)], epi=0)
    with pytest.raises(RuntimeError, match="workspace"):
        call([TL.GemmGroup(**base)], nsplit=3)
    rows = (TL.RowsGroup * 1)(TL.RowsGroup(-1, 4, fake, fake, 4, fake, 4, 0))
    with pytest.raises(RuntimeError, match="negative size"):
        lib.tm_rows_grouped(1, ctypes.cast(rows, ctypes.c_void_p), 0, None)
    casts = (TL.CastGroup * 1)(TL.CastGroup(fake, fake, -5))
    with pytest.raises(RuntimeError, match="negative size"):
        lib.tm_cast_bf16_grouped(1, ctypes.cast(casts, ctypes.c_void_p), None)
    with pytest.raises(RuntimeError, match="negative molecule count"):
        lib.tm_fm_loss(-1, *([None] * 12), None)
    packs = (TL.PackPart * 1)(TL.PackPart(fake, 6, 0))
    with pytest.raises(RuntimeError, match="whole 4-byte words"):
        lib.tm_pack_copy(1, ctypes.cast(packs, ctypes.c_void_p), fake, None)
    packs = (TL.PackPart * 1)(TL.PackPart(None, 8, 0))
    with pytest.raises(RuntimeError, match="whole 4-byte words"):
        lib.tm_pack_copy(1, ctypes.cast(packs, ctypes.c_void_p), fake, None)
    with pytest.raises(RuntimeError, match="bad part list"):
        lib.tm_pack_copy(1, None, fake, None)
    lib.tm_pack_copy(0, None, None, None)
    # nothing to do is not an error
    lib.tm_gemm_bf16_grouped(0, None, 0, 0, 0, 0, 1.0, 1, None)
