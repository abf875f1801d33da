"""ctypes binding of libafqmc_hip.so (the C-ABI declared in include/afqmc_hip.h).

The library is built in-tree by ``__graft_entry__.build()`` (``make -C
miniafqmc_amd/csrc``).  There is no fallback: if the shared library is
missing, importing this module raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libafqmc_hip.so")

STATUS = {
    0: "AFQMC_OK",
    1: "AFQMC_ERR_ARG",
    2: "AFQMC_ERR_SHAPE",
    3: "AFQMC_ERR_ALLOC",
    4: "AFQMC_ERR_HIP",
    5: "AFQMC_ERR_SINGULAR",
    6: "AFQMC_ERR_NONFINITE",
    7: "AFQMC_ERR_STATE",
    8: "AFQMC_ERR_UNSUPPORTED",
}

# bytes of the static device scratch array of afqmc_mixed_density_matrix_noalloc
NOALLOC_SCRATCH_BYTES = 64 << 20

MODES = {"fp64_csr": 0, "fp64_dense": 1, "mixed_fp32": 2, "mixed_bf16": 3}

BUF_W, BUF_WDATA, BUF_G, BUF_VBIAS, BUF_X, BUF_VHS, BUF_HYBRIDW, BUF_EPARTS, BUF_GFULL = range(9)
(OPT_DENSE_EXCHANGE, OPT_ENERGY_PRECISION, OPT_ORTHO_HOUSEHOLDER, OPT_TRANSPOSED, OPT_EXCHANGE_CHUNK,
 OPT_FUSED_PROPAGATOR, OPT_DIRECT_LDS, OPT_TILE, OPT_THREE_PIECE, OPT_THREE_PIECE_STEPS,
 OPT_ONE_PIECE_STEPS, OPT_CSR_DENSE, OPT_COULOMB_THREE_PIECE) = range(13)

# every symbol include/afqmc_hip.h declares
EXPORTED = [
    "afqmc_ctx_create", "afqmc_ctx_destroy", "afqmc_ctx_error", "afqmc_sync", "afqmc_set_option",
    "afqmc_upload_system", "afqmc_get_spvnT", "afqmc_init_walkers", "afqmc_copy",
    "afqmc_mixed_density_matrix", "afqmc_vbias", "afqmc_mixed_density_matrix_noalloc",
    "afqmc_mixed_density_matrix_full", "afqmc_vbias_full",
    "afqmc_fields", "afqmc_vhs",
    "afqmc_propagate", "afqmc_save_overlaps", "afqmc_calculate_overlaps",
    "afqmc_weights", "afqmc_branch", "afqmc_orthogonalize", "afqmc_energy", "afqmc_stage_fields",
    "afqmc_device_fields", "afqmc_substep", "afqmc_get_eshift", "afqmc_set_eshift", "afqmc_run", "afqmc_tune",
    "afqmc_use_graphs", "afqmc_profile", "afqmc_kernel_stats", "afqmc_zgemm",
    "afqmc_group_create", "afqmc_group_destroy", "afqmc_group_substep", "afqmc_group_sync", "afqmc_group_error",
    "afqmc_synth_create", "afqmc_synth_destroy", "afqmc_synth_get", "afqmc_synth_fields",
]


class CsrView(C.Structure):
    _fields_ = [
        ("nrows", C.c_int),
        ("ncols", C.c_int),
        ("nnz", C.c_longlong),
        ("rowptr", C.POINTER(C.c_int)),
        ("col", C.POINTER(C.c_int)),
        ("val", C.POINTER(C.c_double)),
    ]


class SynthParams(C.Structure):
    _fields_ = [
        ("nmo", C.c_int),
        ("naea", C.c_int),
        ("nchol", C.c_int),
        ("density", C.c_double),
        ("sigma_L", C.c_double),
        ("sigma_h", C.c_double),
        ("dt", C.c_double),
        ("seed", C.c_uint64),
        ("with_vakbl", C.c_int),
    ]


class SynthView(C.Structure):
    _fields_ = [
        ("nmo", C.c_int),
        ("naea", C.c_int),
        ("nchol", C.c_int),
        ("trial_alpha", C.POINTER(C.c_double)),
        ("trial_beta", C.POINTER(C.c_double)),
        ("haj", C.POINTER(C.c_double)),
        ("propg1", C.POINTER(C.c_double)),
        ("h1", C.POINTER(C.c_double)),
        ("spvn", CsrView),
        ("vakbl", CsrView),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C miniafqmc_amd/csrc)")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    D = C.POINTER(C.c_double)
    I = C.c_int
    L = C.c_long
    sig = {
        "afqmc_ctx_create": (I, [I, I, I, I, I, C.POINTER(P)]),
        "afqmc_ctx_destroy": (I, [P]),
        "afqmc_ctx_error": (C.c_char_p, [P]),
        "afqmc_sync": (I, [P]),
        "afqmc_set_option": (I, [P, I, C.c_double]),
        "afqmc_upload_system": (I, [P, D, D, D, D, C.POINTER(CsrView), C.POINTER(CsrView), C.c_double, C.c_double]),
        "afqmc_get_spvnT": (I, [P, C.POINTER(C.c_longlong), C.POINTER(C.c_int), C.POINTER(C.c_int), D]),
        "afqmc_init_walkers": (I, [P]),
        "afqmc_copy": (I, [P, I, I, D]),
        "afqmc_mixed_density_matrix": (I, [P]),
        "afqmc_vbias": (I, [P]),
        "afqmc_mixed_density_matrix_noalloc": (I, [I, I, I, P, P, P, P, P]),
        "afqmc_mixed_density_matrix_full": (I, [P]),
        "afqmc_vbias_full": (I, [P]),
        "afqmc_fields": (I, [P, D, L]),
        "afqmc_vhs": (I, [P]),
        "afqmc_propagate": (I, [P]),
        "afqmc_save_overlaps": (I, [P]),
        "afqmc_calculate_overlaps": (I, [P]),
        "afqmc_weights": (I, [P, C.c_double, D]),
        "afqmc_branch": (I, [P, C.c_double, C.POINTER(C.c_int)]),
        "afqmc_orthogonalize": (I, [P]),
        "afqmc_energy": (I, [P, D]),
        "afqmc_stage_fields": (I, [P, D, L]),
        "afqmc_device_fields": (I, [P, C.c_uint64, I]),
        "afqmc_substep": (I, [P, L, I, C.c_double]),
        "afqmc_get_eshift": (I, [P, D]),
        "afqmc_set_eshift": (I, [P, C.c_double]),
        "afqmc_run": (I, [P, L, I, I, I, C.c_double, D]),
        "afqmc_tune": (I, [P, C.c_double]),
        "afqmc_use_graphs": (I, [P, I]),
        "afqmc_group_create": (I, [C.POINTER(C.c_void_p), I, C.POINTER(C.c_void_p)]),
        "afqmc_group_destroy": (I, [P]),
        "afqmc_group_substep": (I, [P, C.c_long, I, C.c_double]),
        "afqmc_group_sync": (I, [P]),
        "afqmc_group_error": (C.c_char_p, [P]),
        "afqmc_profile": (I, [P, I]),
        "afqmc_kernel_stats": (I, [P, I, P, C.POINTER(C.c_longlong), D, C.POINTER(C.c_int)]),
        "afqmc_zgemm": (I, [I, C.c_char, C.c_char, I, I, I, D, D, I, D, I, D, D, I]),
        "afqmc_synth_create": (I, [C.POINTER(SynthParams), C.POINTER(P)]),
        "afqmc_synth_destroy": (I, [P]),
        "afqmc_synth_get": (I, [P, C.POINTER(SynthView)]),
        "afqmc_synth_fields": (I, [C.c_uint64, I, I, I, I, L, D]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


class AfqmcError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__(f"{STATUS.get(code, code)}: {msg}")


def check(code, ctx=None):
    if code != 0:
        msg = ""
        if ctx is not None:
            m = lib.afqmc_ctx_error(ctx)
            msg = m.decode() if m else ""
        raise AfqmcError(code, msg)
