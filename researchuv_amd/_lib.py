"""Loader for libruv_lscm.so (the C ABI in include/ruv_lscm.h).

The library is built in-tree by researchuv_amd/build.py.  There is no CPU
fallback: a missing library raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libruv_lscm.so")

RUV_OK = 0
STATUS = {
    0: "OK",
    -1: "UNAVAILABLE",
    -2: "BAD_DIMS",
    -3: "BREAKDOWN",
    -4: "NOT_CONVERGED",
    -5: "NONFINITE",
    -6: "DEVICE",
    -7: "PRECOND",
    -8: "INVALID",
}


class RuvError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class SolveInfo(C.Structure):
    _fields_ = [
        ("iters", C.c_uint64),
        ("refinements", C.c_uint32),
        ("precond", C.c_int32),
        ("true_rel_resid", C.c_double),
        ("seconds", C.c_double),
    ]


class UnfoldOpts(C.Structure):
    _fields_ = [
        ("max_iter", C.c_int32),
        ("angle_mix", C.c_double),
        ("mix_w", C.c_double),
        ("precond", C.c_int32),
        ("mode", C.c_int32),
        ("min_vars", C.c_int32),
        ("tol", C.c_double),
        ("device_assembly", C.c_int32),
    ]


_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        # torch ships its own copy of the HIP runtime (same SONAME).  Loading torch first makes this library
        # bind to that copy, so a process that also uses torch.cuda has exactly one HIP runtime; loading this
        # library first would leave torch a second runtime that finds no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python researchuv_amd/build.py` (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.ruv_available.restype = C.c_int
        L.ruv_last_error.restype = C.c_char_p
        L.ruv_cg_solve_precond.restype = C.c_int
        L.ruv_cg_solve_precond.argtypes = [P, P, P, C.c_int32, C.c_int64, P, C.c_int64, C.c_uint64, C.c_double,
                                           C.c_int32, P, P]
        L.ruv_cg_solve.restype = C.c_int
        L.ruv_cg_solve.argtypes = [P, P, P, C.c_int32, C.c_int64, P, C.c_int64, C.c_uint64, C.c_double, P]
        L.ruv_cg_solve_mixed.restype = C.c_int
        L.ruv_cg_solve_mixed.argtypes = [P, P, P, C.c_int32, C.c_int64, P, C.c_int64, C.c_uint64, C.c_double,
                                         C.c_int32, P, P]
        L.ruv_apply_ic0.restype = C.c_int
        L.ruv_apply_ic0.argtypes = [P, P, P, C.c_int32, C.c_int64, P, P]
        L.ruv_plan_create.restype = C.c_int
        L.ruv_plan_create.argtypes = [P, P, P, C.c_int32, C.c_int64, C.c_int32, C.c_int32, C.c_int32, P]
        L.ruv_plan_solve_dev.restype = C.c_int
        L.ruv_plan_solve_dev.argtypes = [P, P, P, C.c_uint64, C.c_double, P]
        L.ruv_plan_solve.restype = C.c_int
        L.ruv_plan_solve.argtypes = [P, P, P, C.c_uint64, C.c_double, P]
        L.ruv_plan_destroy.restype = None
        L.ruv_plan_destroy.argtypes = [P]
        L.ruv_plan_update_values.restype = C.c_int
        L.ruv_plan_update_values.argtypes = [P, P]
        L.ruv_ic0_factor.restype = C.c_int
        L.ruv_ic0_factor.argtypes = [P, P, P, C.c_int32, C.c_int64, P, P, P, P]
        for name in ("ruv_unfold_chart",):
            if hasattr(L, name):
                f = getattr(L, name)
                f.restype = C.c_int
        _lib = L
    return _lib


def last_error() -> str:
    return lib().ruv_last_error().decode()


def check(status: int):
    if status != RUV_OK:
        raise RuvError(status, last_error())
