"""ctypes binding of libtnad.so (include/tnad.h).

The library is the product; this module only declares its C signatures.  It
fails loudly when the shared object is missing or does not export a symbol:
there is no CPU fallback behind it.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libtnad.so"

TNAD_OK = 0
ERRORS = {
    -1: "TNAD_ERR_ARG",
    -2: "TNAD_ERR_SHAPE",
    -3: "TNAD_ERR_HIP",
    -4: "TNAD_ERR_NAN",
    -5: "TNAD_ERR_ALLOC",
    -6: "TNAD_ERR_UNSUPPORTED",
}
ERR_ARG, ERR_SHAPE, ERR_HIP, ERR_NAN, ERR_ALLOC, ERR_UNSUPPORTED = -1, -2, -3, -4, -5, -6
PREC = {"fp64": 0, "fp32": 1, "bf16": 2}


class Opts(C.Structure):
    _fields_ = [
        ("precision", C.c_int),
        ("switch_tol", C.c_double),
        ("bwd_precision", C.c_int),
        ("eig_block", C.c_int),
        ("eig_max_iters", C.c_int),
        ("eig_warm", C.c_int),
        ("eig_warm_iters", C.c_int),
        ("eig_tol", C.c_double),
        ("bwd_last", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("iters", C.c_int),
        ("low_prec_iters", C.c_int),
        ("eig_iters_total", C.c_int),
        ("eig_k", C.c_int),
        ("eig_resid_max", C.c_double),
        ("bwd_steps", C.c_int),
        ("fwd_ms", C.c_double),
        ("eig_ms", C.c_double),
        ("bwd_ms", C.c_double),
    ]


P = C.c_void_p
I = C.c_int
D = C.c_double
PI = C.POINTER(C.c_int)
PD = C.POINTER(C.c_double)

# name -> (restype, argtypes); every symbol declared in include/tnad.h
SIGNATURES = {
    "tnad_version": (I, []),
    "tnad_opts_default": (None, [C.POINTER(Opts)]),
    "tnad_create": (I, [I, C.POINTER(P)]),
    "tnad_destroy": (I, [P]),
    "tnad_last_error": (C.c_char_p, [P]),
    "tnad_get_stats": (I, [P, C.POINTER(Stats)]),
    "tnad_indexperm_symmetrize": (I, [P, P, I, I, P]),
    "tnad_double_layer": (I, [P, P, I, I, P, P]),
    "tnad_init_raw": (I, [P, P, I, I, P, P]),
    "tnad_symeig": (I, [P, P, I, I, I, P, P, PI]),
    "tnad_symeig_noalloc": (I, [P, P, I, P, P]),
    "tnad_svd_back_sym": (I, [P, P, I, I, I, P, P]),
    "tnad_ctmrgstep": (I, [P, P, I, I, P, P, C.POINTER(Opts), P, P, P]),
    "tnad_ctmrg": (I, [P, P, I, I, P, P, D, I, C.POINTER(Opts), P, P, P, PI, C.POINTER(P)]),
    "tnad_ctmrg_back": (I, [P, P, P, P, P]),
    "tnad_tape_free": (I, [P]),
    "tnad_expectationvalue": (I, [P, P, I, P, I, I, P, P, P]),
    "tnad_energy": (I, [P, P, P, I, I, I, D, I, C.POINTER(Opts), PD, P, PI]),
    "tnad_trg": (I, [P, P, I, I, I, D, PD, P]),
    "tnad_trg_svd": (I, [P, P, I, I, I, I, I, D, P, P, PI]),
    "tnad_trg_svd_back": (I, [P, P, I, I, I, I, I, D, P, P, I, P]),
    "tnad_energy_z": (I, [P, P, P, I, I, I, D, I, P, P, PI]),
    "tnad_trg_z": (I, [P, P, I, I, I, D, P, P]),
    "tnad_svd_z": (I, [P, P, I, I, P, P, P]),
    "tnad_svd_back_z": (I, [P, P, I, I, P, P, P, P]),
    "tnad_contract": (I, [P, I, C.c_char_p, P, P, C.c_char_p, P, P, C.c_char_p, P, D, D, P]),
    "tnad_time_contract": (I, [P, I, C.c_char_p, P, C.c_char_p, P, C.c_char_p, P, I, PD]),
    "tnad_time_jacobi": (I, [P, I, I, I, PD, PI]),
    "tnad_time_syev": (I, [P, I, I, PD]),
    "tnad_mlp_create": (I, [P, I, I, I, I, C.c_ulonglong, C.POINTER(P)]),
    "tnad_mlp_free": (I, [P]),
    "tnad_mlp_num_params": (I, [P, C.POINTER(C.c_longlong)]),
    "tnad_mlp_get_params": (I, [P, P, P]),
    "tnad_mlp_set_params": (I, [P, P, P]),
    "tnad_mlp_forward": (I, [P, P, P, I, P]),
    "tnad_mlp_loss_grad": (I, [P, P, P, P, I, PD, P]),
    "tnad_mlp_train_step": (I, [P, P, P, P, I, C.c_float, PD]),
    "tnad_bulk_features": (I, [P, P, I, I, P]),
    "tnad_env_features": (I, [P, P, P, I, I, P]),
    "tnad_env_from_features": (I, [P, P, I, I, P, P]),
}

_lib = None


def load():
    """Load libtnad.so once; raise if it is missing or incomplete."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"libtnad.so not found at {LIB_PATH}: build it with "
            "`python tensornetworkad.jl_amd/build_lib.py` (hipcc, gfx950)")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | getattr(os, "RTLD_GLOBAL", 0))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)          # AttributeError if the symbol is not exported
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class TnadError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class DimensionMismatch(TnadError, ValueError):
    """Raised where the reference throws DimensionMismatch (src/ipeps.jl:19-20)."""
