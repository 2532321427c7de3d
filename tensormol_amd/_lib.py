"""ctypes binding of libtmamd.so, the C-ABI HIP library.

Every entry point takes plain device pointers, sizes and a hipStream_t and
returns 0 or a negative status; failures raise RuntimeError carrying
tm_last_error().  There is no fallback: if the library is missing, importing
the engine fails.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtmamd.so")

P = ctypes.c_void_p
I = ctypes.c_int
L = ctypes.c_long
D = ctypes.c_double
F = ctypes.c_float

# name -> argument types (all return int status except where noted)
SIGNATURES = {
    "tm_version": [],
    "tm_zero_i32": [I, P, P],
    "tm_cell_build": [I, P, P, D, D, D, D, D, D, I, I, I, I, P, P, P, P, P, P, P, P, P, P],
    "tm_nl_build": [I, P, P, P, P, P, P, P, P, P, P, P, I, I, I, D, D, D, D, D, D, I, I, D, D, D, I, I, P, P, P,
                    P, P, P, P, P, P],
    "tm_triples": [I, I, I, I, P, P, P, P, P, P, P],
    "tm_pairs": [I, P, P, P, P, P, P, P, P, P, P, P, I, I, I, D, D, D, D, D, D, I, D, D, I, P, P, P, P, P],
    "tm_sym_forward": [I, P, P, P, P, P, P, P, P, P, P, I, I, I, I, F, F, F, F, P, P, P, P, I, P, P, P],
    "tm_sym_adjoint": [I, P, P, P, P, P, P, P, P, P, P, I, I, I, I, F, F, F, F, P, P, P, P, P, P,
                       P, P, I, P, P],
    "tm_sym_jvp": [I, P, P, P, P, P, P, P, P, P, P, I, I, I, I, F, F, F, F, P, P, P, P, P, I, P],
    "tm_force_gather": [I, I, P, P, P, P, D, P, P, P],
    "tm_force_gather_vv": [I, I, P, P, P, P, D, P, P, D, P, P, P, P, P, P],
    "tm_mlp_f32": [I, I, I, P, P, I, P, P, F, P, P, P, P],
    "tm_mol_sum_f32": [I, P, I, P, I, P, I, P, I, P],
    "tm_mol_sum_f64": [I, P, I, P, I, P, I, P, I, P],
    "tm_mol_sums": [I, P, I, I, P, P, P, P, P, P, P],
    "tm_mol_sums_static": [I, P, I, I, P, P, P, P, P, P],
    "tm_charges": [I, P, P, P, P, P, P, P, P, P, P, P, P, P],
    "tm_charge_upstream": [I, P, P, P, P, P],
    "tm_coulomb": [I, P, P, P, P, P, P, P, P, P, P, I, I, I, D, D, D, D, D, D, I, P, D, D, I, I, I, P, P, P, P,
                   P, P, P, P, P],
    "tm_coulomb_table": [P, D, D, I, I, I, P, P, I, P, P],
    "tm_coulomb_tab": [I, P, P, P, P, I, I, I, D, D, D, D, D, D, I, D, I, P, I, D, D, P, P, P, P,
                           P, P, P],
    "tm_vv_drift": [I, D, P, P, P, P],
    "tm_vv_kick": [I, D, P, P, P, P, P, P],
    "tm_vv_kick_drift": [I, D, P, P, P, P, P, P, P],
    "tm_nh_drift": [I, D, P, P, P, P, P, P, P, P, P],
    "tm_nh_eta": [I, P, D, P, P, P, P, P],
    "tm_nh_kick": [I, D, P, P, P, P, P, P, P, P, P],
    "tm_fire_dots": [I, P, P, P, P],
    "tm_fire_update": [I, I, P, I, P, P, P, P, P, P, P, D, I, D, D, D, D, D, P],
    "tm_mlp_mfma": [I, I, I, I, P, P, I, P, P, P, F, I, I, P, P, P],
    "tm_gemm_bf16": [I, I, I, P, I, I, P, I, I, P, I, I, P, P, P, F, I, P, P],
    "tm_gemm_bf16_grouped": [I, P, I, I, I, I, F, I, P],
    "tm_train_elem": [I, I, I, I, F, P, P, P, P, P, P, P, P],
    "tm_rows": [I, I, I, P, P, I, P, I, P],
    "tm_rows_grouped": [I, P, I, P],
    "tm_cast_bf16_grouped": [I, P, P],
    "tm_fm_loss": [I, P, P, P, P, P, P, P, P, P, P, P, P, P],
    "tm_pack_copy": [I, P, P, P],
    "tm_mlp_mfma_pack_info": [I, I, I, I, P],
}


class _Lib:
    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise ImportError(f"{path} not built: run `python -m tensormol_amd.build` "
                              "(the HIP library is required; there is no CPU fallback)")
        self.path = path
        self.dll = ctypes.CDLL(path)
        self.dll.tm_last_error.restype = ctypes.c_char_p
        self.dll.tm_last_error.argtypes = []
        self.fns = {}
        for name, args in SIGNATURES.items():
            fn = getattr(self.dll, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = ctypes.c_int
            self.fns[name] = fn

    def has(self, name):
        return name in self.fns

    def __getattr__(self, name):
        fns = self.__dict__.get("fns", {})
        if name not in fns:
            raise AttributeError(name)
        fn = fns[name]

        err = self.dll.tm_last_error

        def call(*args):
            st = fn(*args)
            if st < 0:
                msg = err().decode(errors="replace")
                raise RuntimeError(f"{name} failed ({st}): {msg}")
            return st
        # kept as an instance attribute: later lookups skip __getattr__
        self.__dict__[name] = call
        return call


class GemmGroup(ctypes.Structure):
    """tm_gemm_group: one GEMM of a tm_gemm_bf16_grouped call."""
    _fields_ = [("M", I), ("N", I), ("K", I), ("A", P), ("lda", I), ("B", P), ("ldb", I),
                ("C", P), ("ldc", I), ("bias", P), ("s", P), ("aux", P), ("work", P),
                ("x1", P), ("x2", P), ("out2", P), ("crow", P)]


class RowsGroup(ctypes.Structure):
    """tm_rows_group: one row gather / scatter of a tm_rows_grouped call."""
    _fields_ = [("M", I), ("cols", I), ("idx", P), ("inp", P), ("ldi", I), ("out", P), ("ldo", I),
                ("out_bf16", I)]


class PackPart(ctypes.Structure):
    """tm_pack_part: one array of a tm_pack_copy call."""
    _fields_ = [("src", P), ("bytes", L), ("offset", L)]


class CastGroup(ctypes.Structure):
    """tm_cast_group: one array of a tm_cast_bf16_grouped call."""
    _fields_ = [("inp", P), ("out", P), ("n", L)]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _Lib()
    return _LIB


def ptr(t):
    """Device (or host) pointer of a torch tensor, None for None."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def fptr(arr):
    """ctypes pointer to a contiguous float32 numpy array (host parameters)."""
    return arr.ctypes.data_as(ctypes.c_void_p)
