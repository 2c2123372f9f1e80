"""Python mirror of the researchuv-gpu solver interface (crates/researchuv-gpu/src/solver.rs, precond.rs),
backed by the HIP C ABI.

Same names, argument meaning and error behaviour as the reference: `GpuSolver.global_()` is `None` without a
device, and `cg_solve_precond` returns `None` on every failure (breakdown, non-convergence, non-finite,
bad dimensions) exactly where the Rust returned `Option::None`.  The reason is kept in `last_error`.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from . import _lib


class Precond(enum.IntEnum):
    """precond.rs:20-47"""

    NONE = 0
    JACOBI = 1
    IC0 = 2
    IC0_COLOR = 3
    AMG = 4
    AMG_SA = 5

    @staticmethod
    def from_env_name(name: Optional[str]) -> "Precond":
        """RESEARCHUV_GPU_PRECOND parsing (lscm.rs:492-499); anything else is Jacobi."""
        return {
            "amgsa": Precond.AMG_SA,
            "amg": Precond.AMG,
            "ic0color": Precond.IC0_COLOR,
            "ic0": Precond.IC0,
            "none": Precond.NONE,
        }.get(name or "", Precond.JACOBI)


@dataclass
class CsrMatrix:
    """solver.rs:14-44"""

    vals: np.ndarray
    cols: np.ndarray
    row_ptr: np.ndarray

    def __post_init__(self):
        self.vals = np.ascontiguousarray(self.vals, np.float64)
        self.cols = np.ascontiguousarray(self.cols, np.int32)
        self.row_ptr = np.ascontiguousarray(self.row_ptr, np.int32)

    def rows(self) -> int:
        return max(len(self.row_ptr) - 1, 0)

    def row(self, i):
        lo, hi = self.row_ptr[i], self.row_ptr[i + 1]
        return zip(self.cols[lo:hi].tolist(), self.vals[lo:hi].tolist())

    def row_get(self, i, j) -> float:
        for c, v in self.row(i):
            if c == j:
                return v
        return 0.0


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class GpuSolver:
    """solver.rs:49-616 mirror.  One process-wide instance (`global_`)."""

    _global = None
    _probed = False

    def __init__(self):
        self._L = _lib.lib()
        self.last_error = ""

    @staticmethod
    def new() -> Optional["GpuSolver"]:
        """solver.rs:122-201: None when no device is usable."""
        L = _lib.lib()
        return GpuSolver() if L.ruv_available() else None

    @classmethod
    def global_(cls) -> Optional["GpuSolver"]:
        """solver.rs:104-113"""
        if not cls._probed:
            cls._global = cls.new()
            cls._probed = True
        return cls._global

    def _args(self, a: CsrMatrix, b):
        b = np.ascontiguousarray(b, np.float64)
        return b, (_p(a.vals), _p(a.cols), _p(a.row_ptr), a.rows(), len(a.vals))

    def cg_solve_precond(self, a: CsrMatrix, b, max_iter: int, tol: float,
                         precond: Precond) -> Optional[Tuple[np.ndarray, int]]:
        """solver.rs:284-565"""
        b, csr = self._args(a, b)
        x = np.zeros(a.rows())
        it = C.c_uint64(0)
        rc = self._L.ruv_cg_solve_precond(*csr, _p(b), len(b), max_iter, tol, int(precond), _p(x), C.byref(it))
        if rc != 0:
            self.last_error = f"{_lib.STATUS.get(rc, rc)}: {_lib.last_error()}"
            return None
        return x, int(it.value)

    def cg_solve(self, a: CsrMatrix, b, max_iter: int, tol: float) -> Optional[np.ndarray]:
        """solver.rs:274-277"""
        r = self.cg_solve_precond(a, b, max_iter, tol, Precond.NONE)
        return None if r is None else r[0]

    def cg_solve_mixed(self, a: CsrMatrix, b, max_iter: int, tol: float, precond: Precond):
        """FP32 PCG inside FP64 iterative refinement; tol on the true residual.  Returns (x, info) or None."""
        b, csr = self._args(a, b)
        x = np.zeros(a.rows())
        info = _lib.SolveInfo()
        rc = self._L.ruv_cg_solve_mixed(*csr, _p(b), len(b), max_iter, tol, int(precond), _p(x), C.byref(info))
        if rc != 0:
            self.last_error = f"{_lib.STATUS.get(rc, rc)}: {_lib.last_error()}"
            return None
        return x, info

    def apply_ic0_for_test(self, a: CsrMatrix, r) -> Optional[np.ndarray]:
        """solver.rs:205-270"""
        r = np.ascontiguousarray(r, np.float64)
        if len(r) != a.rows():
            return None
        z = np.zeros_like(r)
        rc = self._L.ruv_apply_ic0(_p(a.vals), _p(a.cols), _p(a.row_ptr), a.rows(), len(a.vals), _p(r), _p(z))
        if rc != 0:
            self.last_error = f"{_lib.STATUS.get(rc, rc)}: {_lib.last_error()}"
            return None
        return z


class Plan:
    """A device-resident system (upload and preconditioner setup once; solve many times)."""

    def __init__(self, a: CsrMatrix, rhs_per_row: int, precond: Precond, mixed: bool):
        self._L = _lib.lib()
        self.h = C.c_void_p()
        self.rows = a.rows()
        self.R = rhs_per_row
        _lib.check(self._L.ruv_plan_create(_p(a.vals), _p(a.cols), _p(a.row_ptr), a.rows(), len(a.vals),
                                           rhs_per_row, int(precond), 1 if mixed else 0, C.byref(self.h)))

    def solve(self, b, max_iter: int, tol: float):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b)
        info = _lib.SolveInfo()
        _lib.check(self._L.ruv_plan_solve(self.h, _p(b), _p(x), max_iter, tol, C.byref(info)))
        return x, info

    def solve_dev(self, b_ptr: int, x_ptr: int, max_iter: int, tol: float):
        """b_ptr, x_ptr: device addresses (e.g. torch tensor.data_ptr())."""
        info = _lib.SolveInfo()
        _lib.check(self._L.ruv_plan_solve_dev(self.h, C.c_void_p(b_ptr), C.c_void_p(x_ptr), max_iter, tol,
                                              C.byref(info)))
        return info

    def update_values(self, vals):
        """New values for the plan's pattern (the CSR order of creation): operator and Jacobi / IC(0)
        preconditioner refreshed on the device."""
        vals = np.ascontiguousarray(vals, np.float64)
        _lib.check(self._L.ruv_plan_update_values(self.h, _p(vals)))

    def close(self):
        if self.h:
            self._L.ruv_plan_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ic0_factor(a: CsrMatrix) -> dict:
    """The IC(0) factor of `a` as the device factorisation computes it (factor_ic0, precond.rs:145-219): L's rows,
    strict-lower entries ascending and the diagonal last."""
    L = _lib.lib()
    nl = C.c_int64(0)
    args = (_p(a.vals), _p(a.cols), _p(a.row_ptr), a.rows(), len(a.vals))
    _lib.check(L.ruv_ic0_factor(*args, C.byref(nl), None, None, None))
    ptr = np.zeros(a.rows() + 1, np.int32)
    cols = np.zeros(max(nl.value, 1), np.int32)
    vals = np.zeros(max(nl.value, 1))
    _lib.check(L.ruv_ic0_factor(*args, C.byref(nl), _p(ptr), _p(cols), _p(vals)))
    return dict(l_ptr=ptr, l_cols=cols[:nl.value], l_vals=vals[:nl.value])


class AmgHost:
    """The plan path's host AMG hierarchy (amg.rs:127-455), level by level (an inspection hook for tests)."""

    def __init__(self, a: CsrMatrix, smoothed: bool):
        L = _lib.lib()
        P = C.c_void_p
        L.ruv_amg_host_build.restype = C.c_int
        L.ruv_amg_host_build.argtypes = [P, P, P, C.c_int32, C.c_int64, C.c_int32, P]
        L.ruv_amg_host_levels.restype = C.c_int32
        L.ruv_amg_host_levels.argtypes = [P]
        L.ruv_amg_host_level.restype = C.c_int
        L.ruv_amg_host_level.argtypes = [P, C.c_int32, P, P, P, P, P, P, P, P]
        L.ruv_amg_host_destroy.restype = None
        L.ruv_amg_host_destroy.argtypes = [P]
        self._L = L
        self.h = C.c_void_p()
        _lib.check(L.ruv_amg_host_build(_p(a.vals), _p(a.cols), _p(a.row_ptr), a.rows(), len(a.vals),
                                        1 if smoothed else 0, C.byref(self.h)))

    def __len__(self):
        return int(self._L.ruv_amg_host_levels(self.h))

    def level(self, k: int) -> dict:
        """op / p as scipy CSR, agg, coarse rows, factored flag."""
        import scipy.sparse as sp

        sz = np.zeros(5, np.int64)
        nul = None
        _lib.check(self._L.ruv_amg_host_level(self.h, k, _p(sz), nul, nul, nul, nul, nul, nul, nul))
        n, nnz, pnz, nc, fac = (int(v) for v in sz)
        ov, oc, op = np.zeros(nnz), np.zeros(nnz, np.int32), np.zeros(n + 1, np.int32)
        pv, pc, pp = np.zeros(pnz), np.zeros(pnz, np.int32), np.zeros(n + 1, np.int32)
        agg = np.zeros(n if nc else 0, np.int32)
        _lib.check(self._L.ruv_amg_host_level(self.h, k, _p(sz), _p(ov), _p(oc), _p(op), _p(pv), _p(pc), _p(pp),
                                              _p(agg) if nc else None))
        out = dict(op=sp.csr_matrix((ov, oc, op), shape=(n, n)), n_coarse=nc, factored=bool(fac), agg=agg)
        out["p"] = sp.csr_matrix((pv, pc, pp), shape=(n, nc)) if nc else None
        return out

    def close(self):
        if self.h:
            self._L.ruv_amg_host_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def amg_aggregate(a: CsrMatrix, max_members: int):
    """aggregate (amg.rs:254-291) -> (agg, n_coarse)"""
    L = _lib.lib()
    L.ruv_amg_aggregate.restype = C.c_int
    L.ruv_amg_aggregate.argtypes = [C.c_void_p] * 3 + [C.c_int32, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p]
    agg = np.zeros(a.rows(), np.int32)
    nc = C.c_int32(0)
    _lib.check(L.ruv_amg_aggregate(_p(a.vals), _p(a.cols), _p(a.row_ptr), a.rows(), len(a.vals), max_members,
                                   _p(agg), C.byref(nc)))
    return agg, int(nc.value)


def _csr_call(fn, args):
    """Two-call protocol of the host CSR hooks: sizes first, then the arrays -> scipy CSR."""
    import scipy.sparse as sp

    sz = np.zeros(2, np.int64)
    _lib.check(fn(*args, _p(sz), None, None, None))
    n, nnz = int(sz[0]), int(sz[1])
    v, c, p = np.zeros(nnz), np.zeros(nnz, np.int32), np.zeros(n + 1, np.int32)
    _lib.check(fn(*args, _p(sz), _p(v), _p(c), _p(p)))
    return v, c, p, n


def amg_prolongator(a: CsrMatrix, agg, n_coarse: int, smoothed: bool):
    """prolongator (amg.rs:308-362) -> scipy CSR (rows x n_coarse)"""
    import scipy.sparse as sp

    L = _lib.lib()
    L.ruv_amg_prolongator.restype = C.c_int
    L.ruv_amg_prolongator.argtypes = [C.c_void_p] * 3 + [C.c_int32, C.c_int64, C.c_void_p, C.c_int32, C.c_int32] + \
        [C.c_void_p] * 4
    agg = np.ascontiguousarray(agg, np.int32)
    v, c, p, n = _csr_call(L.ruv_amg_prolongator, (_p(a.vals), _p(a.cols), _p(a.row_ptr), a.rows(), len(a.vals),
                                                   _p(agg), n_coarse, 1 if smoothed else 0))
    return sp.csr_matrix((v, c, p), shape=(n, n_coarse))


def amg_galerkin(a: CsrMatrix, p_csr, n_coarse: int):
    """galerkin_p (amg.rs:384-431) -> scipy CSR (n_coarse x n_coarse)"""
    import scipy.sparse as sp

    L = _lib.lib()
    L.ruv_amg_galerkin.restype = C.c_int
    L.ruv_amg_galerkin.argtypes = [C.c_void_p] * 3 + [C.c_int32, C.c_int64] + [C.c_void_p] * 3 + [C.c_int32] + \
        [C.c_void_p] * 4
    pv = np.ascontiguousarray(p_csr.data, np.float64)
    pc = np.ascontiguousarray(p_csr.indices, np.int32)
    pp = np.ascontiguousarray(p_csr.indptr, np.int32)
    v, c, p, n = _csr_call(L.ruv_amg_galerkin, (_p(a.vals), _p(a.cols), _p(a.row_ptr), a.rows(), len(a.vals),
                                                _p(pv), _p(pc), _p(pp), n_coarse))
    return sp.csr_matrix((v, c, p), shape=(n, n))
