"""ctypes binding of libkmx.so (the C ABI declared in include/kmx.h).

The library is the only compute path: if it is missing or fails to load, every
entry point raises -- there is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libkmx.so")

KMX_OK = 0
KMX_E_ARG = -1
KMX_E_SHAPE = -2
KMX_E_HIP = -3
KMX_E_NOMEM = -4
KMX_E_UNSUPPORTED = -5

PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16 = 3

KRON_SCRATCH_ELEMS = 1 << 20
ARNOLDI_MAX_K = 127


def arnoldi_work(n):
    """KMX_ARNOLDI_WORK(n): doubles of caller-owned scratch for kmx_arnoldi_step."""
    return int(n) + 32896


PLAN_FP32 = 1
PLAN_ADJOINT = 2
PLAN_LEVEL_BLOCKS = 4

BUF_X, BUF_Y, BUF_W0, BUF_W1, BUF_FACTOR = 0, 1, 2, 3, 4

_CODES = {
    KMX_E_ARG: "KMX_E_ARG",
    KMX_E_SHAPE: "KMX_E_SHAPE",
    KMX_E_HIP: "KMX_E_HIP",
    KMX_E_NOMEM: "KMX_E_NOMEM",
    KMX_E_UNSUPPORTED: "KMX_E_UNSUPPORTED",
}


class KmxError(RuntimeError):
    """An entry point returned a negative status; .code holds it."""

    def __init__(self, code, msg):
        super().__init__(f"{_CODES.get(code, code)}: {msg}")
        self.code = code


class PlanDesc(C.Structure):
    _fields_ = [
        ("ndim", C.c_int),
        ("deg", C.c_int),
        ("num_elem", C.c_int),
        ("num_terms", C.c_int),
        ("mat_rows", C.POINTER(C.c_int)),
        ("mats", C.POINTER(C.c_void_p)),
        ("elem_idx1d", C.POINTER(C.c_int)),
        ("row_ptr", C.POINTER(C.c_int64)),
        ("cols", C.POINTER(C.c_int)),
        ("index_base", C.c_int),
        ("flags", C.c_uint),
        ("identity", C.POINTER(C.c_ubyte)),
    ]


class SweepInfo(C.Structure):
    _fields_ = [
        ("available", C.c_int),
        ("separable", C.c_int),
        ("fp64", C.c_int),
        ("identity_dev", C.c_double),
        ("drop_tol", C.c_double),
        ("flops", C.c_double),
        ("mfma_flops", C.c_double),
        ("kept_blocks", C.c_int64),
        ("total_blocks", C.c_int64),
        ("operator_values", C.c_int64),
        ("box_values", C.c_int64),
    ]


PTERM_MASS, PTERM_DIV, PTERM_GRAD = 0, 1, 2


class PtermDesc(C.Structure):
    _fields_ = [
        ("type", C.c_int),
        ("LF", C.c_double),
        ("ibcl", C.c_char),
        ("ibcr", C.c_char),
        ("g_quad", C.POINTER(C.c_double)),
        ("g_face", C.POINTER(C.c_double)),
        ("lhs_quad", C.POINTER(C.c_double)),
    ]


class GemmDesc(C.Structure):
    _fields_ = [
        ("stage", C.c_int),
        ("transA", C.c_char),
        ("transB", C.c_char),
        ("m", C.c_int),
        ("n", C.c_int),
        ("k", C.c_int),
        ("lda", C.c_int),
        ("ldb", C.c_int),
        ("ldc", C.c_int),
        ("alpha", C.c_double),
        ("beta", C.c_double),
        ("bufA", C.c_int),
        ("bufB", C.c_int),
        ("bufC", C.c_int),
        ("offA", C.c_int64),
        ("offB", C.c_int64),
        ("offC", C.c_int64),
    ]


# every symbol include/kmx.h declares, with its prototype
PROTOTYPES = {
    "kmx_plan_create": (C.c_int, [C.POINTER(PlanDesc), C.POINTER(C.c_void_p)]),
    "kmx_plan_update_term": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "kmx_plan_set_term_scale": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "kmx_plan_destroy": (C.c_int, [C.c_void_p]),
    "kmx_plan_autotune": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kmx_plan_get_variant": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "kmx_plan_set_sweep": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "kmx_plan_get_sweep": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_double)]),
    "kmx_plan_sweep_info": (C.c_int, [C.c_void_p, C.POINTER(SweepInfo)]),
    "kmx_realspace_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "kmx_realspace_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kmx_realspace_destroy": (C.c_int, [C.c_void_p]),
    "kmx_plan_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int), C.POINTER(C.c_int64)]),
    "kmx_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "kmx_apply_transpose": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "kmx_plan_assemble": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "kmx_apply_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "kmx_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]),
    "kmx_batch_apply": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int, C.c_void_p]),
    "kmx_batch_destroy": (C.c_int, [C.c_void_p]),
    "kmx_batch_set_sweep": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "kmx_batch_get_sweep": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
    "kmx_batch_autotune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_void_p]),
    "kmx_batch_rk3": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_double), C.c_int, C.c_int,
                                C.c_void_p]),
    "kmx_kronmult": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_void_p), C.c_void_p,
                               C.c_int, C.c_void_p, C.c_void_p]),
    "kmx_arnoldi_step": (C.c_int, [C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                   C.c_void_p, C.c_void_p, C.c_void_p]),
    "kmx_kronmult_noalloc": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_void_p),
                                       C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "kmx_gemm_vbatched": (C.c_int, [C.c_char, C.c_char, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.c_double, C.POINTER(C.c_void_p), C.POINTER(C.c_int), C.POINTER(C.c_void_p),
                                    C.POINTER(C.c_int), C.c_double, C.POINTER(C.c_void_p), C.POINTER(C.c_int), C.c_int,
                                    C.c_void_p]),
    "kmx_kronmult_batch_plan": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int,
                                          C.POINTER(GemmDesc), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int64)]),
    "kmx_fmwt_matrix": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kmx_quad_points": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_int), C.c_void_p]),
    "kmx_sd_term_matrix": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_int,
                                     C.POINTER(PtermDesc), C.c_void_p, C.c_void_p]),
    "kmx_last_error": (C.c_char_p, []),
    "kmx_sync": (C.c_int, [C.c_void_p]),
    "kmx_version": (C.c_char_p, []),
}

_lib = None


def lib():
    """Load libkmx.so once (RuntimeError if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libkmx.so not built ({LIB_PATH}); run python dg-sparsegrid_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != KMX_OK:
        raise KmxError(rc, lib().kmx_last_error().decode(errors="replace"))
    return rc
